// MI355X (gfx950) finite-volume compressible RHS: device kernels and the C ABI (include/ablate_fv.h).
//
// One RHS evaluation (FiniteVolumeSolver::ComputeRHSFunction, finiteVolumeSolver.cpp:233-291, plus the ghost
// fill and aux update the TS callback runs before it, timeStepper.cpp:331-392) is, per call:
//   1. k_ghost_fill        boundary faces: ghost-cell state from the boundary condition (essentialGhost.cpp:33-51)
//   2. k_gradients         one lane per interior cell: least-squares gradients of the conserved state
//                          (CellInterpolant::ComputeFieldGradients, cellInterpolant.cpp:330-413) and their limiter
//                          (:415-487), plus the unlimited aux (T, velocity) gradients the face-interpolant
//                          viscous flux averages (LeastSquaresAverage, stencils/leastSquaresAverage.cpp:5-155)
//   3. k_species_gradients same for densityYi / Yi when species are transported
//   4. k_flux              one lane per interior cell: for each of its faces (fixed face-number order), project
//                          both sides to the face (ProjectToFace, cellInterpolant.cpp:890-950), evaluate the face
//                          flux exactly as the reference orients it (left = support[0]) and gather -F/V_L or +F/V_R
//                          into the cell's residual.  Every face flux is evaluated by both of its cells instead of
//                          being stored: no float atomics, no face-flux array, bitwise deterministic.
#include <hip/hip_runtime.h>

#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "fv_adjoint.hpp"
#include "fv_adjoint2.hpp"
#include "fv_device.hpp"
#include "fv_internal.hpp"
#include "fv_layout.hpp"

namespace abl {

namespace {
thread_local std::string g_last_error;
}
void set_global_error(const std::string& msg) { g_last_error = msg; }
void resolve_noalloc_scratch();

constexpr int kBlock = 256;
constexpr int kTileMax = 512;           // largest tile (workgroup) of the tile strategy
constexpr int kTileDefault = 256;       // largest tile the default choice takes
constexpr int kTileLdsMax = 160 * 1024;  // LDS of one workgroup (gfx950: the whole 160 KB of a CU)
constexpr int kTileLdsDefault = 64 * 1024;  // LDS budget of the default choice (two or more workgroups per CU)
constexpr int kCuLdsBytes = 160 * 1024;  // LDS of one gfx950 compute unit

// ============================================================================================ kernels

__global__ __launch_bounds__(kBlock) void k_ghost_fill(int64_t nb, const int32_t* __restrict__ bface,
                                                       const int32_t* __restrict__ bslot, const int32_t* __restrict__ fcell,
                                                       int64_t F, int64_t N, int nc, const int32_t* __restrict__ slot_kind,
                                                       const double* __restrict__ slot_state, double* __restrict__ U) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nb) return;
    const int32_t f = bface[i], s = bslot[i];
    const int32_t interior = fcell[f], ghost = fcell[F + f];
    const int kind = slot_kind[s];
    if (kind == ABL_FV_BC_GHOST) return;  // caller-filled: the ghost rows are inputs
    for (int p = 0; p < nc; p++) {
        const double ui = U[(int64_t)p * N + interior];
        double v = ui;  // OUTFLOW: zero-gradient extrapolation
        if (kind == ABL_FV_BC_INFLOW) v = slot_state[s * nc + p];
        if (kind == kBcInflowAtFace) v = 2.0 * slot_state[s * nc + p] - ui;  // essentialGhost.cpp:17-20
        U[(int64_t)p * N + ghost] = v;
    }
}

// Aux values of a cell: temperature (PerfectGas TemperatureFunction, perfectGas.cpp:76-86) and velocity.
template <int DIM>
__device__ inline void aux_of(const double* u, const Phys& ph, double* a) {
    const double irho = 1.0 / u[kRho];
    double ke = 0.0;
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        const double v = u[kRhoU + d] * irho;
        a[1 + d] = v;
        ke += v * v;
    }
    a[0] = (u[kRhoE] * irho - 0.5 * ke) * ph.icv;
}

// Limiter families resolved at compile time.
enum : int { kLimNone = 0, kLimPetsc = 1, kLimBJ = 2 };

// Cell record of the FP64 form, written by the gradient pass for every interior cell and read by the face
// fluxes: words [0, NE) the state, [G0, G0 + NE*DIM) the limited euler gradients (p * DIM + d), [T0, T0 + NT) the
// viscous terms.  Records are stored in blocks of 64 consecutive cells, word-pair-major inside a block: words
// (2q, 2q + 1) of cell c form the 16-byte pair at double2 index rec_pair(c) + 64 q.  The 64 lanes of a wave (64
// consecutive cells) store each pair with one coalesced 1 KB store, and the faces of a wave, whose cells cluster,
// read each pair from a few cache lines with one 16-byte load (a per-cell record made every load of the wave touch
// up to 64 lines; single words per load took twice the load instructions).
template <int DIM>
struct CellRec {
    static constexpr int NE = DIM + 2;
    static constexpr int G0 = NE;
    static constexpr int T0 = NE + NE * DIM;
    static constexpr int END = T0 + ViscTerms<DIM>::NT;
    static constexpr int NP = (END + 1) / 2;  // word pairs per cell
    static constexpr int RS = 2 * NP;         // words per cell (the last may be padding)
    static constexpr int BLOCK = 64;
};
template <int DIM>
__device__ inline int64_t rec_pair(int64_t c) {
    return (c >> 6) * (int64_t)(CellRec<DIM>::BLOCK * CellRec<DIM>::NP) + (c & 63);
}
// Word i of cell c's record.
template <int DIM>
__device__ inline double rec_word(const double* __restrict__ rec, int64_t c, int i) {
    return rec[2 * (rec_pair<DIM>(c) + (int64_t)CellRec<DIM>::BLOCK * (i >> 1)) + (i & 1)];
}

// FP64 species gradients: per species and cell 2 DIM words, [0, DIM) the limited densityYi gradient and
// [DIM, 2 DIM) the Yi aux gradient, stored as DIM word pairs; pair j of species sp of cell c is the double2 at
// (sp DIM + j) N + c.  Lanes of consecutive cells move each pair with one coalesced 16-byte access, and a face's
// right cell (scattered) costs DIM cache lines per species instead of 2 DIM.
template <int DIM>
__device__ inline int64_t sgp_word(int64_t N, int sp, int64_t c, int w) {
    return (((int64_t)sp * DIM + (w >> 1)) * N + c) * 2 + (w & 1);
}

// Cell record of the mixed form, in 64-bit words: [0, U0/2) pairs of FP32 values (the limited gradients at float
// [0, NE*DIM), the viscous terms at float [TF, TF + NT)), then the NE FP64 state words.  Stored in blocks of 64
// cells in word pairs like the FP64 record: words (2q, 2q + 1) of cell c at 16-byte index recf_pair(c) + 64 q.
template <int DIM>
struct CellRecF {
    static constexpr int NE = DIM + 2;
    static constexpr int TF = NE * DIM;
    static constexpr int U0 = ((TF + ViscTerms<DIM>::NT) + 1) & ~1;  // float offset of the state (even)
    static constexpr int RSW = U0 / 2 + NE;                          // 64-bit words in use per cell
    static constexpr int NP = (RSW + 1) / 2;                          // word pairs per cell
    static constexpr int RS = 4 * NP;                                 // floats per cell (with padding)
};
template <int DIM>
__device__ inline int64_t recf_pair(int64_t c) {
    return (c >> 6) * (int64_t)(64 * CellRecF<DIM>::NP) + (c & 63);
}

template <int DIM, bool RECON, bool VISC>
struct RecF {
    double u[DIM + 2];
    float g[RECON ? (DIM + 2) * DIM : 1];
    float tq[VISC ? ViscTerms<DIM>::NT : 1];
};
// Loads a mixed-form cell record (FP64 state, FP32 gradients and viscous terms).
template <int DIM, bool RECON, bool VISC>
__device__ inline void load_rec_f(const float* __restrict__ rec, int64_t c, RecF<DIM, RECON, VISC>& r) {
    using CR = CellRecF<DIM>;
    constexpr int NE = DIM + 2;
    const ulonglong2* __restrict__ b = reinterpret_cast<const ulonglong2*>(rec) + recf_pair<DIM>(c);
    constexpr int NF = VISC ? CR::TF + ViscTerms<DIM>::NT : (RECON ? CR::TF : 0);
    constexpr int NFW = (NF + 1) / 2;  // words holding the FP32 values needed
    uint64_t wd[2 * CR::NP];
#pragma unroll
    for (int q = 0; q < CR::NP; q++) {
        const bool need = 2 * q < NFW || (2 * q + 1 >= CR::U0 / 2 && 2 * q < CR::U0 / 2 + NE);
        const ulonglong2 v = need ? b[64 * q] : make_ulonglong2(0ull, 0ull);
        wd[2 * q] = v.x;
        wd[2 * q + 1] = v.y;
    }
    float t[NF > 0 ? ((NF + 1) & ~1) : 2];
#pragma unroll
    for (int i = 0; i < NFW; i++) {
        t[2 * i] = __uint_as_float((uint32_t)wd[i]);
        t[2 * i + 1] = __uint_as_float((uint32_t)(wd[i] >> 32));
    }
    if (RECON)
#pragma unroll
        for (int i = 0; i < NE * DIM; i++) r.g[i] = t[i];
    if (VISC)
#pragma unroll
        for (int q = 0; q < ViscTerms<DIM>::NT; q++) r.tq[q] = t[CR::TF + q];
#pragma unroll
    for (int p = 0; p < NE; p++) r.u[p] = __longlong_as_double((long long)wd[CR::U0 / 2 + p]);
}

// Gradients of the euler field (limited) and of the aux T/velocity (unlimited), one interior cell per lane.
// G is the storage/arithmetic type of weights and gradients (double: FP64 form, float: mixed form); state
// differences are always taken in FP64 before rounding to G.  Each lane writes its cell's record: CellRec (FP64
// form) or CellRecF (mixed form), word by word into the 64-cell blocks.
template <typename G, int DIM, bool RECON, bool VISC, int LIM, int KM>
__device__ __forceinline__ void gradients_body(MeshView<G> m, Phys ph, const double* __restrict__ U,
                                               G* __restrict__ gradC, G* __restrict__ gradA, void* __restrict__ rec) {
    constexpr int NE = DIM + 2;
    constexpr int NA = DIM + 1;
    const int64_t c_lane = xcd_cell_index();
    const bool active = c_lane < m.NI;
    const int64_t c = active ? c_lane : m.NI - 1;  // lanes past the end shadow the last cell and store nothing
    const int64_t N = m.N, NI = m.NI;
    const G* __restrict__ wa_arr = m.swa ? m.swa : m.sw;
    const bool own_wa = m.swa != nullptr;  // meshes without boundary faces share one weight set
    double uc[NE], ac[NA];
#pragma unroll
    for (int p = 0; p < NE; p++) uc[p] = U[p * N + c];
    if (VISC) aux_of<DIM>(uc, ph, ac);
    G g[NE][DIM], ga[NA][DIM];
    // Barth-Jespersen bounds as the extreme neighbour differences u_n - u_c, each rounded to G once: rounding is
    // monotone, so they equal G(max u_n - u_c) and G(min u_n - u_c) bit for bit, in half the registers of FP64 bounds
    G dmin[NE], dmax[NE];
#pragma unroll
    for (int p = 0; p < NE; p++) {
        dmin[p] = G(0);
        dmax[p] = G(0);
#pragma unroll
        for (int d = 0; d < DIM; d++) g[p][d] = G(0);
    }
#pragma unroll
    for (int q = 0; q < NA; q++)
#pragma unroll
        for (int d = 0; d < DIM; d++) ga[q][d] = G(0);

#pragma unroll
    for (int k = 0; k < KM; k++) {
        const int32_t e = m.snbr[k * NI + c];
        const bool valid = e >= 0;
        const int64_t n = valid ? (int64_t)e : c;
        const bool bnd = n >= NI;
        double un[NE];
#pragma unroll
        for (int p = 0; p < NE; p++) un[p] = U[p * N + n];
        G w[DIM], wa[DIM];
#pragma unroll
        for (int d = 0; d < DIM; d++)
            if (RECON) w[d] = m.sw[(k * DIM + d) * NI + c];
        if (VISC) {
            if (RECON && !own_wa) {  // the aux set is the cell set: reuse the words just loaded
#pragma unroll
                for (int d = 0; d < DIM; d++) wa[d] = w[d];
            } else {
#pragma unroll
                for (int d = 0; d < DIM; d++) wa[d] = wa_arr[(k * DIM + d) * NI + c];
            }
        }
        if (!valid) continue;
        if (RECON) {
#pragma unroll
            for (int p = 0; p < NE; p++) {
                const G delta = G(un[p] - uc[p]);
                if (LIM == kLimBJ) {
                    dmin[p] = fmin(dmin[p], delta);
                    dmax[p] = fmax(dmax[p], delta);
                }
                // cellInterpolant.cpp:405-411 with delta = u_R - u_L: the left cell adds w0.delta and the right cell
                // subtracts w1.delta, i.e. both add w.(u_n - u_c) with their own side's weight (bitwise the same)
                if (!bnd)
#pragma unroll
                    for (int d = 0; d < DIM; d++) g[p][d] += w[d] * delta;
            }
        }
        if (VISC) {
            double an[NA];
            aux_of<DIM>(un, ph, an);
#pragma unroll
            for (int q = 0; q < NA; q++) {
                const G delta = G(an[q] - ac[q]);
#pragma unroll
                for (int d = 0; d < DIM; d++) ga[q][d] += wa[d] * delta;
            }
        }
    }

    if (RECON && LIM != kLimNone) {
        G phi[NE];
        if (LIM == kLimBJ) {
            // phi = min(1, (u_max - u_c)/max_f D_f+, (u_min - u_c)/min_f D_f-), D_f = grad.(x_f - x_c): the per-face
            // minimum collapsed onto the extreme projections (division is monotone).
            G dpos[NE], dneg[NE];
#pragma unroll
            for (int p = 0; p < NE; p++) {
                dpos[p] = G(0);
                dneg[p] = G(0);
            }
#pragma unroll
            for (int k = 0; k < KM; k++) {
                const bool valid = m.snbr[k * NI + c] >= 0;
                G x[DIM];
#pragma unroll
                for (int d = 0; d < DIM; d++) x[d] = m.slim[(k * DIM + d) * NI + c];
                if (!valid) continue;
#pragma unroll
                for (int p = 0; p < NE; p++) {
                    G D = G(0);
#pragma unroll
                    for (int d = 0; d < DIM; d++) D += g[p][d] * x[d];
                    dpos[p] = fmax(dpos[p], D);
                    dneg[p] = fmin(dneg[p], D);
                }
            }
#pragma unroll
            for (int p = 0; p < NE; p++) {
                G v = G(1);
                if (dpos[p] > G(0)) v = fmin(v, dmax[p] / dpos[p]);
                if (dneg[p] < G(0)) v = fmin(v, dmin[p] / dneg[p]);
                phi[p] = v;
            }
        } else {
            // DMPlexApplyLimiter_Internal (cellInterpolant.cpp:282-328): f = 1/2 (u_n - u_c) / (grad.(x_n - x_c))
#pragma unroll
            for (int p = 0; p < NE; p++) phi[p] = sizeof(G) == 8 ? G(DBL_MAX) : G(FLT_MAX);
#pragma unroll
            for (int k = 0; k < KM; k++) {
                const int32_t e = m.snbr[k * NI + c];
                const int64_t n = e >= 0 ? (int64_t)e : c;
                G v[DIM];  // x_n - x_c
#pragma unroll
                for (int d = 0; d < DIM; d++) v[d] = m.slim[(k * DIM + d) * NI + c];
                if (e < 0) continue;
#pragma unroll
                for (int p = 0; p < NE; p++) {
                    G denom = G(0);
#pragma unroll
                    for (int d = 0; d < DIM; d++) denom += g[p][d] * v[d];
                    const G flim = G(0.5) * G(U[p * N + n] - uc[p]) / denom;
                    phi[p] = pmin(phi[p], limiter_phi<G>(ph.limiter, flim));
                }
            }
        }
        bool cancel = false;
#pragma unroll
        for (int p = 0; p < NE; p++)
            if (phi[p] == G(0))
#pragma unroll
                for (int d = 0; d < DIM; d++)
                    if (double(fabs(g[p][d])) > ph.max_lim_grad) cancel = true;
#pragma unroll
        for (int p = 0; p < NE; p++)
#pragma unroll
            for (int d = 0; d < DIM; d++) g[p][d] *= cancel ? G(0) : phi[p];
    }
    G tq[ViscTerms<DIM>::NT];
    if (VISC) stress_heat<G, DIM>(G(ph.mu), G(ph.k), ga, tq);
    if (sizeof(G) == 8) {
        using CR = CellRec<DIM>;
        if (active) {
            double w[CR::RS];
#pragma unroll
            for (int p = 0; p < NE; p++) {
                w[p] = uc[p];
#pragma unroll
                for (int d = 0; d < DIM; d++) w[CR::G0 + p * DIM + d] = RECON ? double(g[p][d]) : 0.0;
            }
#pragma unroll
            for (int q = 0; q < ViscTerms<DIM>::NT; q++) w[CR::T0 + q] = VISC ? double(tq[q]) : 0.0;
#pragma unroll
            for (int i = CR::END; i < CR::RS; i++) w[i] = 0.0;
            double2* __restrict__ r = reinterpret_cast<double2*>(rec) + rec_pair<DIM>(c_lane);
#pragma unroll
            for (int q = 0; q < CR::NP; q++) r[CR::BLOCK * q] = make_double2(w[2 * q], w[2 * q + 1]);
        }
    } else {
        // FP32 record: float pairs packed little-endian into 64-bit words, then the FP64 state
        using CR = CellRecF<DIM>;
        float t[CR::U0];
#pragma unroll
        for (int i = 0; i < CR::U0; i++) t[i] = 0.f;
#pragma unroll
        for (int p = 0; p < NE; p++)
#pragma unroll
            for (int d = 0; d < DIM; d++)
                if (RECON) t[p * DIM + d] = float(g[p][d]);
#pragma unroll
        for (int q = 0; q < ViscTerms<DIM>::NT; q++)
            if (VISC) t[CR::TF + q] = float(tq[q]);
        if (active) {
            uint64_t wd[2 * CR::NP];
#pragma unroll
            for (int i = 0; i < CR::U0 / 2; i++)
                wd[i] = (uint64_t)__float_as_uint(t[2 * i]) | ((uint64_t)__float_as_uint(t[2 * i + 1]) << 32);
#pragma unroll
            for (int p = 0; p < NE; p++) wd[CR::U0 / 2 + p] = (uint64_t)__double_as_longlong(uc[p]);
#pragma unroll
            for (int i = CR::RSW; i < 2 * CR::NP; i++) wd[i] = 0;
            ulonglong2* __restrict__ w = reinterpret_cast<ulonglong2*>(rec) + recf_pair<DIM>(c_lane);
#pragma unroll
            for (int q = 0; q < CR::NP; q++) w[64 * q] = make_ulonglong2(wd[2 * q], wd[2 * q + 1]);
        }
    }
}

template <typename G, int DIM, bool RECON, bool VISC, int LIM, int KM>
__global__ __launch_bounds__(kBlock) void k_gradients(MeshView<G> m, Phys ph, const double* __restrict__ U,
                                                      G* __restrict__ gradC, G* __restrict__ gradA,
                                                      void* __restrict__ rec) {
    gradients_body<G, DIM, RECON, VISC, LIM, KM>(m, ph, U, gradC, gradA, rec);
}
// The same kernel held to 128 VGPRs (four waves per SIMD instead of three: more loads in flight).  Picked for the
// FP64 3-D forms without the PETSc-family limiter, which fit without spilling (that limiter's would spill).
template <typename G, int DIM, bool RECON, bool VISC, int LIM, int KM>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4))) void k_gradients_w4(
    MeshView<G> m, Phys ph, const double* __restrict__ U, G* __restrict__ gradC, G* __restrict__ gradA,
    void* __restrict__ rec) {
    gradients_body<G, DIM, RECON, VISC, LIM, KM>(m, ph, U, gradC, gradA, rec);
}

// densityYi gradients (limited, one field) into gS [ns*DIM][N] and Yi aux gradients (unlimited) into gY [ns*DIM][N];
// FP64 arithmetic, stored as G (double: FP64 form, float: mixed form).  The cell's slot geometry (neighbours,
// weights, limiter vectors, neighbour densities) is loaded once for all species.
template <typename G, int DIM, int KM>
__device__ __forceinline__ void species_body(MeshView<G> m, Phys ph, const double* __restrict__ U,
                                             G* __restrict__ gradC, G* __restrict__ gradA) {
    constexpr int NE = DIM + 2;
    const int64_t c = xcd_cell_index();
    if (c >= m.NI) return;
    const int64_t N = m.N, NI = m.NI;
    const bool recon = ph.recon != 0;
    const bool aux = ph.sdiff != 0;
    const bool bj = ph.limiter == ABL_FV_LIMITER_BARTH_JESPERSEN;
    const G* __restrict__ wa_arr = m.swa ? m.swa : m.sw;
    const double irho_c = 1.0 / U[kRho * N + c];
    // arithmetic of the gradient sums and the limiter: FP64 in the FP64 form; FP32 in the mixed form (which stores
    // these gradients in FP32), on FP64 state differences rounded once
    using A = G;
    int64_t nb[KM];
    bool valid[KM], inner[KM];
    G w[KM][DIM], wa[KM][DIM], xl[KM][DIM];
    double irn[KM];
#pragma unroll
    for (int k = 0; k < KM; k++) {
        const int32_t e = m.snbr[k * NI + c];
        valid[k] = e >= 0;
        nb[k] = valid[k] ? (int64_t)e : c;
        inner[k] = valid[k] && nb[k] < NI;
#pragma unroll
        for (int d = 0; d < DIM; d++) {
            w[k][d] = recon ? m.sw[(k * DIM + d) * NI + c] : G(0);
            wa[k][d] = aux ? wa_arr[(k * DIM + d) * NI + c] : G(0);
            xl[k][d] = (recon && m.slim) ? m.slim[(k * DIM + d) * NI + c] : G(0);
        }
        irn[k] = aux ? 1.0 / U[kRho * N + nb[k]] : 0.0;
    }
    bool cancel = false;
    for (int sp = 0; sp < ph.ns; sp++) {
        const int64_t comp = NE + sp;
        const double uc = U[comp * N + c];
        const double yc = uc * irho_c;
        double un[KM];
#pragma unroll
        for (int k = 0; k < KM; k++) un[k] = U[comp * N + nb[k]];
        A g[DIM], ga[DIM];
        double umin = uc, umax = uc;
#pragma unroll
        for (int d = 0; d < DIM; d++) {
            g[d] = A(0);
            ga[d] = A(0);
        }
#pragma unroll
        for (int k = 0; k < KM; k++) {
            if (!valid[k]) continue;
            if (recon) {
                umin = fmin(umin, un[k]);
                umax = fmax(umax, un[k]);
                if (inner[k]) {
                    const A delta = A(un[k] - uc);  // both sides: + w.(u_n - u_c), see k_gradients
#pragma unroll
                    for (int d = 0; d < DIM; d++) g[d] += w[k][d] * delta;
                }
            }
            if (aux) {
                const A delta = A(un[k] * irn[k] - yc);
#pragma unroll
                for (int d = 0; d < DIM; d++) ga[d] += wa[k][d] * delta;
            }
        }
        if (recon) {
            A phi = sizeof(A) == 8 ? A(DBL_MAX) : A(FLT_MAX);
            if (bj) {
                A dpos = A(0), dneg = A(0);
#pragma unroll
                for (int k = 0; k < KM; k++) {
                    if (!valid[k]) continue;
                    A D = A(0);
#pragma unroll
                    for (int d = 0; d < DIM; d++) D += g[d] * xl[k][d];
                    dpos = fmax(dpos, D);
                    dneg = fmin(dneg, D);
                }
                phi = A(1);
                if (dpos > A(0)) phi = fmin(phi, A(umax - uc) / dpos);
                if (dneg < A(0)) phi = fmin(phi, A(umin - uc) / dneg);
            } else if (ph.limiter != ABL_FV_LIMITER_NONE) {
#pragma unroll
                for (int k = 0; k < KM; k++) {
                    if (!valid[k]) continue;
                    A denom = A(0);
#pragma unroll
                    for (int d = 0; d < DIM; d++) denom += g[d] * xl[k][d];
                    const A flim = A(0.5) * A(un[k] - uc) / denom;
                    phi = pmin(phi, limiter_phi<A>(ph.limiter, flim));
                }
            } else {
                phi = A(1);
            }
            if (phi == A(0))
#pragma unroll
                for (int d = 0; d < DIM; d++)
                    if (fabs(double(g[d])) > ph.max_lim_grad) cancel = true;
#pragma unroll
            for (int d = 0; d < DIM; d++) g[d] *= phi;  // stored with the aux gradient below
        }
        {
            using V2 = typename Vec2T<G>::type;
            G v[2 * DIM];
#pragma unroll
            for (int d = 0; d < DIM; d++) {
                v[d] = recon ? G(g[d]) : G(0);
                v[DIM + d] = aux ? G(ga[d]) : G(0);
            }
            V2* __restrict__ o = reinterpret_cast<V2*>(gradC);
#pragma unroll
            for (int j = 0; j < DIM; j++) {
                V2 t;
                t.x = v[2 * j];
                t.y = v[2 * j + 1];
                o[((int64_t)sp * DIM + j) * N + c] = t;
            }
        }
    }
    if (cancel)  // the maxLimGrad rule zeroes the whole densityYi field of the cell (cellInterpolant.cpp:459-481)
        for (int sp = 0; sp < ph.ns; sp++)
#pragma unroll
            for (int d = 0; d < DIM; d++) gradC[sgp_word<DIM>(N, sp, c, d)] = G(0);
    (void)gradA;
}

template <typename G, int DIM, int KM>
__global__ __launch_bounds__(kBlock) void k_species_gradients(MeshView<G> m, Phys ph, const double* __restrict__ U,
                                                              G* __restrict__ gradC, G* __restrict__ gradA) {
    species_body<G, DIM, KM>(m, ph, U, gradC, gradA);
}
// The same kernel held to three waves per SIMD (168 VGPRs), picked for 4-slot cells (triangles, tetrahedra).
template <typename G, int DIM, int KM>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(3))) void k_species_gradients_w3(
    MeshView<G> m, Phys ph, const double* __restrict__ U, G* __restrict__ gradC, G* __restrict__ gradA) {
    species_body<G, DIM, KM>(m, ph, U, gradC, gradA);
}

// Everything a face flux needs from one side: the state projected to the face (ProjectToFace,
// cellInterpolant.cpp:890-950), the cell density, temperature and velocity (face-interpolant values) and the cell's
// viscous terms.
template <int DIM, bool VISC>
struct SideData {
    static constexpr int NE = DIM + 2;
    double uf[NE];
    double rho;
    double T;
    double vel[DIM];
    double tq[VISC ? ViscTerms<DIM>::NT : 1];
};

// Primitive face-interpolant values of a side from its cell state: density, velocity and internal energy
// (scaled to T by the caller's cv only where needed).
template <int DIM, bool VISC>
__device__ inline void side_cell_values(const double* u, SideData<DIM, VISC>& sd) {
    sd.rho = u[kRho];
    if (VISC) {
        const double irho = 1.0 / u[kRho];
        double ke = 0.0;
#pragma unroll
        for (int d = 0; d < DIM; d++) {
            sd.vel[d] = u[kRhoU + d] * irho;
            ke += sd.vel[d] * sd.vel[d];
        }
        sd.T = u[kRhoE] * irho - 0.5 * ke;
    }
}

// Loads side data of cell c for a face whose centroid is at x_c + dx: from the cell record when the gradient pass
// ran (reconstruction or diffusion), else from the state.  Ghost cells have zero gradients and read the state;
// their viscous terms are filled by ghost_visc.  (Each branch keeps its own arrays so that they stay in
// registers.)
template <int DIM, bool RECON, bool VISC>
__device__ inline void load_side(const double* __restrict__ U, const double* __restrict__ rec, int64_t N, int64_t c,
                                 bool ghost, const double* dx, SideData<DIM, VISC>& sd) {
    using CR = CellRec<DIM>;
    constexpr int NE = DIM + 2;
    if ((RECON || VISC) && !ghost) {
        constexpr int NL = VISC ? CR::END : CR::T0;
        double r[2 * ((NL + 1) / 2)];
        const double2* __restrict__ rb = reinterpret_cast<const double2*>(rec) + rec_pair<DIM>(c);
#pragma unroll
        for (int q = 0; q < (NL + 1) / 2; q++) {
            // without reconstruction the pairs holding only gradient words are not read
            const bool need = RECON || 2 * q < CR::G0 || 2 * q + 1 >= CR::T0;
            const double2 v = need ? rb[CR::BLOCK * q] : make_double2(0.0, 0.0);
            r[2 * q] = v.x;
            r[2 * q + 1] = v.y;
        }
#pragma unroll
        for (int p2 = 0; p2 < NE; p2++) {
            double a = 0.0;
            if (RECON)
#pragma unroll
                for (int d = 0; d < DIM; d++) a += r[CR::G0 + p2 * DIM + d] * dx[d];
            sd.uf[p2] = r[p2] + a;
        }
        if (VISC)
#pragma unroll
            for (int q = 0; q < ViscTerms<DIM>::NT; q++) sd.tq[q] = r[CR::T0 + q];
        side_cell_values<DIM, VISC>(r, sd);
    } else {
        double u[NE];
#pragma unroll
        for (int p2 = 0; p2 < NE; p2++) {
            u[p2] = U[p2 * N + c];
            sd.uf[p2] = u[p2];
        }
        if (VISC)
#pragma unroll
            for (int q = 0; q < ViscTerms<DIM>::NT; q++) sd.tq[q] = 0.0;
        side_cell_values<DIM, VISC>(u, sd);
    }
}

// Viscous terms of a ghost cell: its one-neighbour least-squares gradient w (a_interior - a_ghost)
// (the LeastSquaresAverage stencil of a ghost cell, stencils/leastSquaresAverage.cpp:69-155).
template <int DIM, bool VISC>
__device__ inline void ghost_visc(const Phys& ph, const double* w, const SideData<DIM, VISC>& in, SideData<DIM, VISC>& gh) {
    double gA[DIM + 1][DIM];
    const double dT = (in.T - gh.T) * ph.icv;
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        gA[0][d] = w[d] * dT;
#pragma unroll
        for (int q = 0; q < DIM; q++) gA[1 + q][d] = w[d] * (in.vel[q] - gh.vel[q]);
    }
    stress_heat<double, DIM>(ph.mu, ph.k, gA, gh.tq);
}

// The flux through one face oriented left -> right (ComputeFluxSourceTerms + the continuous face loop,
// cellInterpolant.cpp:561-631 and faceInterpolant.cpp:356-435): the advection flux of the projected states and
// the viscous flux from face-averaged viscous terms.  Also returns the coefficients of the species fluxes:
// F_sp = cl * rhoY_L + cr * rhoY_R + cd * n.(gradY_L + gradY_R).
template <int DIM, bool VISC>
__device__ inline void face_flux(const Phys& ph, const double* nhat, double area, const SideData<DIM, VISC>& L,
                                 const SideData<DIM, VISC>& R, double* flux, double& cl, double& cr, double& cd) {
    constexpr int NE = DIM + 2;
#pragma unroll
    for (int p = 0; p < NE; p++) flux[p] = 0.0;
    cl = 0.0;
    cr = 0.0;
    cd = 0.0;
    if (ph.flux == ABL_FV_FLUX_RUSANOV) {
        double unL, unR, lam;
        rusanov_flux<double, DIM>(L.uf, R.uf, nhat, area, ph, flux, unL, unR, lam);
        cl = area * (0.5 * unL + 0.5 * lam);
        cr = area * (0.5 * unR - 0.5 * lam);
    } else if (ph.flux != ABL_FV_FLUX_NONE) {
        double mA, rl, rr;
        int dir;
        advection_flux<double, DIM>(ph.flux, L.uf, R.uf, nhat, area, ph, flux, mA, dir, rl, rr);
        species_coefficients<double>(dir, ph.compact != 0, mA, rl, rr, cl, cr);
    }
    if (VISC) {
        double vel[DIM], tq[ViscTerms<DIM>::NT], normal[DIM];
#pragma unroll
        for (int d = 0; d < DIM; d++) {
            vel[d] = 0.5 * L.vel[d] + 0.5 * R.vel[d];
            normal[d] = area * nhat[d];
        }
#pragma unroll
        for (int q = 0; q < ViscTerms<DIM>::NT; q++) tq[q] = 0.5 * L.tq[q] + 0.5 * R.tq[q];
        const double rho_f = 0.5 * L.rho + 0.5 * R.rho;
        double diff = ph.diff;
        if (ph.suth) {
            // transport properties at the face temperature (the face-interpolated aux T of DiffusionFlux,
            // navierStokesTransport.cpp:466-468): tau scales with mu, the heat flux with k = mu cp / Pr, and the
            // species flux with D = mu / (rho Sc) (sutherland.cpp:9-64)
            const double T_f = (0.5 * L.T + 0.5 * R.T) * ph.icv;
            const double mu_f = sutherland_mu(T_f);
#pragma unroll
            for (int q = 0; q < ViscTerms<DIM>::NTAU; q++) tq[q] *= mu_f;
#pragma unroll
            for (int d = 0; d < DIM; d++) tq[ViscTerms<DIM>::NTAU + d] *= mu_f * ph.cp / kSuthPr;
            diff *= mu_f / rho_f / kSuthSc;
        }
        viscous_flux_add<double, DIM>(normal, vel, tq, flux);
        // species diffusion: -D rho_face n.(1/2 (gradY_L + gradY_R)), rho_face = 1/2 (rho_L + rho_R)
        cd = -diff * rho_f * 0.5 * area;
    }
}

// Species fluxes of one face (SpeciesTransport AdvectionFlux / DiffusionSpeciesFlux, speciesTransport.cpp:196-341)
// from the coefficients of face_flux: sum over species of a per-species scalar expression.
template <int DIM, bool RECON, bool VISC>
__device__ inline double species_face_flux(const MeshView<double>& m, const double* __restrict__ U,
                                           const double* __restrict__ gradC, const double* __restrict__ gradA,
                                           int64_t comp, int sp, int64_t f, int64_t L, int64_t R, bool ghost,
                                           const double (*x)[DIM], const double* nhat, double cl, double cr, double cd,
                                           const double* wghost) {
    const int64_t N = m.N;
    double yL = U[comp * N + L], yR = U[comp * N + R];
    const double rL = U[kRho * N + L], rR = U[kRho * N + R];
    // both sides' gradient words (sgp_word pairs); a pair is read only if it holds a word this face uses
    double vL[2 * DIM], vR[2 * DIM];
    const double2* __restrict__ sg = reinterpret_cast<const double2*>(gradC);
#pragma unroll
    for (int j = 0; j < DIM; j++) {
        const bool need = (RECON && 2 * j < DIM) || (VISC && 2 * j + 1 >= DIM);
        const double2 a = need ? sg[((int64_t)sp * DIM + j) * N + L] : make_double2(0.0, 0.0);
        const double2 b = (need && !ghost) ? sg[((int64_t)sp * DIM + j) * N + R] : make_double2(0.0, 0.0);
        vL[2 * j] = a.x;
        vL[2 * j + 1] = a.y;
        vR[2 * j] = b.x;
        vR[2 * j + 1] = b.y;
    }
    double flux = 0.0;
    if (RECON) {
        double pl = 0.0, pr = 0.0;
#pragma unroll
        for (int d = 0; d < DIM; d++) {
            pl += vL[d] * x[0][d];
            if (!ghost) pr += vR[d] * x[1][d];
        }
        flux = cl * (yL + pl) + cr * (yR + pr);
    } else {
        flux = cl * yL + cr * yR;
    }
    if (VISC) {
        double dot = 0.0;
#pragma unroll
        for (int d = 0; d < DIM; d++) {
            const double gl = vL[DIM + d];
            const double gr = ghost ? wghost[d] * (yL / rL - yR / rR) : vR[DIM + d];
            dot += nhat[d] * (gl + gr);
        }
        flux += cd * dot;
    }
    (void)gradA;
    return flux;
}

// Strategy 1 (recompute): face fluxes gathered per cell.  One lane per interior cell; both cells of a face
// evaluate its flux with the reference orientation, so the gathered residual is conservative and
// order-deterministic, with no face-flux storage.
template <int DIM, bool RECON, bool VISC, bool SPECIES>
__global__ __launch_bounds__(kBlock) void k_flux(MeshView<double> m, Phys ph, const double* __restrict__ U,
                                                 const double* __restrict__ gradC, const double* __restrict__ gradA,
                                                 const double* __restrict__ rec, double* __restrict__ R,
                                                 int accumulate) {
    constexpr int NE = DIM + 2;
    using SD = SideData<DIM, VISC>;
    const int64_t c = xcd_cell_index();
    if (c >= m.NI) return;
    const int64_t N = m.N, F = m.F, NI = m.NI;
    const double* __restrict__ wa_arr = m.fwa;  // present whenever ghosts meet diffusion
    double acc[NE];
#pragma unroll
    for (int p = 0; p < NE; p++) acc[p] = 0.0;
    for (int k = 0; k < m.K; k++) {
        const int32_t e = m.ell[k * NI + c];
        if (e < 0) break;
        const int64_t f = e >> 1;
        const int s = e & 1;
        const int64_t L = m.fcell[f], Rc = m.fcell[F + f];
        const bool ghost = Rc >= NI;  // a ghost is always the right cell
        double nh[3], area, x[2][DIM];
        load_face_na(m.fna, f, nh, area);
        load_face_pair<DIM>(m.fdx, f, x);
        SD sl, sr;
        load_side<DIM, RECON, VISC>(U, rec, N, L, false, x[0], sl);
        load_side<DIM, RECON, VISC>(U, rec, N, Rc, ghost, x[1], sr);
        if (VISC && ghost) {
            double w[2][DIM];
            load_face_pair<DIM>(wa_arr, f, w);
            ghost_visc<DIM, VISC>(ph, w[1], sl, sr);
        }
        double flux[NE], cl, cr, cd;
        face_flux<DIM, VISC>(ph, nh, area, sl, sr, flux, cl, cr, cd);
        const double sign = s == 0 ? -1.0 : 1.0;
#pragma unroll
        for (int p = 0; p < NE; p++) acc[p] += sign * flux[p];
    }
    const double iv = m.inv_vol[c];
#pragma unroll
    for (int p = 0; p < NE; p++) {
        const double r = acc[p] * iv;
        R[p * N + c] = accumulate ? R[p * N + c] + r : r;
    }
    if (!SPECIES) return;
    // species: each face's coefficients are re-evaluated per species pass (register-light)
    for (int sp = 0; sp < ph.ns; sp++) {
        const int64_t comp = NE + sp;
        double a = 0.0;
        for (int k = 0; k < m.K; k++) {
            const int32_t e = m.ell[k * NI + c];
            if (e < 0) break;
            const int64_t f = e >> 1;
            const int s = e & 1;
            const int64_t L = m.fcell[f], Rc = m.fcell[F + f];
            const bool ghost = Rc >= NI;
            double nh[3], area, x[2][DIM], w[2][DIM];
            load_face_na(m.fna, f, nh, area);
            load_face_pair<DIM>(m.fdx, f, x);
            if (VISC && ghost) load_face_pair<DIM>(wa_arr, f, w);
            SD sl, sr;
            load_side<DIM, RECON, VISC>(U, rec, N, L, false, x[0], sl);
            load_side<DIM, RECON, VISC>(U, rec, N, Rc, ghost, x[1], sr);
            if (VISC && ghost) ghost_visc<DIM, VISC>(ph, w[1], sl, sr);
            double flux[NE], cl, cr, cd;
            face_flux<DIM, VISC>(ph, nh, area, sl, sr, flux, cl, cr, cd);
            const double fs = species_face_flux<DIM, RECON, VISC>(m, U, gradC, gradA, comp, sp, f, L, Rc, ghost, x, nh,
                                                                   cl, cr, cd, w[1]);
            a += (s == 0 ? -1.0 : 1.0) * fs;
        }
        const double r = a * iv;
        R[comp * N + c] = accumulate ? R[comp * N + c] + r : r;
    }
}

// The flux through face f with the reference orientation: euler components into fe[0, NE), species into
// fs[sp * fs_stride] (shared by the face-flux and tile strategies, which therefore agree bitwise).
// Species flux sp goes to fs[sp * fs_stride], or with PAIRS to fs[(sp / 2) * fs_stride + sp % 2] (species pairs
// interleaved: the face-flux strategy's rows, written and read by the gather two species per 16-byte access; fs
// 16-byte aligned).
template <int DIM, bool RECON, bool VISC, bool SPECIES, bool PAIRS = false>
__device__ inline void eval_face(const MeshView<double>& m, const Phys& ph, const double* __restrict__ U,
                                 const double* __restrict__ gradC, const double* __restrict__ gradA,
                                 const double* __restrict__ rec, int64_t f, double* fe, double* fs,
                                 int64_t fs_stride) {
    using SD = SideData<DIM, VISC>;
    const int64_t N = m.N, F = m.F, NI = m.NI;
    const int64_t L = m.fcell[f], Rc = m.fcell[F + f];
    const bool ghost = Rc >= NI;
    double nh[3], area, x[2][DIM], w[2][DIM];
    load_face_na(m.fna, f, nh, area);
    load_face_pair<DIM>(m.fdx, f, x);
    SD sl, sr;
    load_side<DIM, RECON, VISC>(U, rec, N, L, false, x[0], sl);
    load_side<DIM, RECON, VISC>(U, rec, N, Rc, ghost, x[1], sr);
    if (VISC && ghost) {
        load_face_pair<DIM>(m.fwa, f, w);
        ghost_visc<DIM, VISC>(ph, w[1], sl, sr);
    }
    double cl, cr, cd;
    face_flux<DIM, VISC>(ph, nh, area, sl, sr, fe, cl, cr, cd);
    if (!SPECIES) return;
    if constexpr (PAIRS) {
        // two species per 16-byte store
        for (int q = 0; 2 * q < ph.ns; q++) {
            const int sp = 2 * q;
            const double a = species_face_flux<DIM, RECON, VISC>(m, U, gradC, gradA, DIM + 2 + sp, sp, f, L, Rc, ghost,
                                                                 x, nh, cl, cr, cd, w[1]);
            const double b = sp + 1 < ph.ns ? species_face_flux<DIM, RECON, VISC>(m, U, gradC, gradA, DIM + 3 + sp,
                                                                                 sp + 1, f, L, Rc, ghost, x, nh, cl,
                                                                                 cr, cd, w[1])
                                            : 0.0;
            *reinterpret_cast<double2*>(fs + q * fs_stride) = make_double2(a, b);
        }
    } else {
        for (int sp = 0; sp < ph.ns; sp++)
            fs[sp * fs_stride] = species_face_flux<DIM, RECON, VISC>(m, U, gradC, gradA, DIM + 2 + sp, sp, f, L, Rc,
                                                                     ghost, x, nh, cl, cr, cd, w[1]);
    }
}

// Strategy 2 (face flux + gather), pass 1: one lane per face, flux evaluated once and stored: euler components as a
// padded record [F][FS] (one contiguous read per face in the gather), species as interleaved pairs [ns/2][F][2]
// (coalesced 16-byte stores; species inside the record would make every store of a wave touch 64 cache lines,
// measured 2.5x slower).
template <int DIM, bool RECON, bool VISC, bool SPECIES>
__device__ __forceinline__ void face_flux_body(MeshView<double> m, Phys ph, const double* __restrict__ U,
                                               const double* __restrict__ gradC, const double* __restrict__ gradA,
                                               const double* __restrict__ rec, double* __restrict__ Fe,
                                               double* __restrict__ Fs) {
    constexpr int NE = DIM + 2;
    constexpr int FS = (NE + 1) & ~1;
    const int64_t f = xcd_cell_index();
    if (f >= m.F) return;
    double flux[FS];
    eval_face<DIM, RECON, VISC, SPECIES, true>(m, ph, U, gradC, gradA, rec, f, flux, Fs + 2 * f, 2 * m.F);
    if (FS > NE) flux[FS - 1] = 0.0;
    double2* out = reinterpret_cast<double2*>(Fe + FS * f);
#pragma unroll
    for (int i = 0; i < FS / 2; i++) out[i] = make_double2(flux[2 * i], flux[2 * i + 1]);
}
template <int DIM, bool RECON, bool VISC, bool SPECIES>
__global__ __launch_bounds__(kBlock) void k_face_flux(MeshView<double> m, Phys ph, const double* __restrict__ U,
                                                      const double* __restrict__ gradC, const double* __restrict__ gradA,
                                                      const double* __restrict__ rec, double* __restrict__ Fe,
                                                      double* __restrict__ Fs) {
    face_flux_body<DIM, RECON, VISC, SPECIES>(m, ph, U, gradC, gradA, rec, Fe, Fs);
}


// Strategy 3 (tile): one workgroup per tile of T consecutive interior cells.  Its lanes evaluate every face of the
// tile once into LDS (faces shared with a neighbouring tile are evaluated by both), then each lane gathers its
// cell's faces from LDS in face-number order.  No face-flux round trip through HBM; the face fluxes and the
// gather order are those of strategy 2, so the residual is bitwise the same.
template <int DIM, bool RECON, bool VISC, bool SPECIES, int KM>
__global__ __launch_bounds__(kTileMax) void k_tile_flux(MeshView<double> m, Phys ph, const double* __restrict__ U,
                                                        const double* __restrict__ gradC,
                                                        const double* __restrict__ gradA,
                                                        const double* __restrict__ rec,
                                                        const int32_t* __restrict__ tstart,
                                                        const int32_t* __restrict__ tfaces,
                                                        const int32_t* __restrict__ tslot, int nfm,
                                                        double* __restrict__ R, int accumulate) {
    constexpr int NE = DIM + 2;
    extern __shared__ double lds[];
    double* le = lds;              // [NE][nfm]
    double* ls = lds + NE * nfm;   // [ns][nfm]
    const int64_t t = xcd_block_index();
    const int32_t f0 = tstart[t], nf = tstart[t + 1] - f0;
    // the face number of a lane's next face is loaded one round ahead
    int32_t fnext = (int)threadIdx.x < nf ? tfaces[f0 + threadIdx.x] : 0;
    for (int i = threadIdx.x; i < nf; i += blockDim.x) {
        double flux[NE];
        const int32_t fcur = fnext;
        if (i + (int)blockDim.x < nf) fnext = tfaces[f0 + i + blockDim.x];
        eval_face<DIM, RECON, VISC, SPECIES>(m, ph, U, gradC, gradA, rec, (int64_t)fcur, flux, ls + i, nfm);
#pragma unroll
        for (int p = 0; p < NE; p++) le[p * nfm + i] = flux[p];
    }
    __syncthreads();
    const int64_t c = t * blockDim.x + threadIdx.x;
    if (c >= m.NI) return;
    const int64_t N = m.N, NI = m.NI;
    int li[KM];
    double sg[KM];
#pragma unroll
    for (int k = 0; k < KM; k++) {
        const int32_t e = tslot[k * NI + c];
        li[k] = e >= 0 ? (e >> 1) : 0;
        sg[k] = e < 0 ? 0.0 : ((e & 1) ? 1.0 : -1.0);
    }
    const double iv = m.inv_vol[c];
#pragma unroll
    for (int p = 0; p < NE; p++) {
        double acc = 0.0;
#pragma unroll
        for (int k = 0; k < KM; k++)
            if (sg[k] != 0.0) acc += sg[k] * le[p * nfm + li[k]];
        const double r = acc * iv;
        R[p * N + c] = accumulate ? R[p * N + c] + r : r;
    }
    if (!SPECIES) return;
    for (int sp = 0; sp < ph.ns; sp++) {
        double a = 0.0;
#pragma unroll
        for (int k = 0; k < KM; k++)
            if (sg[k] != 0.0) a += sg[k] * ls[sp * nfm + li[k]];
        const int64_t comp = NE + sp;
        const double r = a * iv;
        R[comp * N + c] = accumulate ? R[comp * N + c] + r : r;
    }
}

// Strategy 2, pass 2: each interior cell gathers its faces' stored fluxes in face-number order.
template <int DIM, bool SPECIES, int KM>
__global__ __launch_bounds__(kBlock) void k_gather(MeshView<double> m, int ns, const double* __restrict__ Fe,
                                                   const double* __restrict__ Fs, double* __restrict__ R, int accumulate) {
    constexpr int NE = DIM + 2;
    constexpr int FS = (NE + 1) & ~1;
    const int64_t c = xcd_cell_index();
    if (c >= m.NI) return;
    const int64_t N = m.N, F = m.F, NI = m.NI;
    int64_t fk[KM];
    double sg[KM];
#pragma unroll
    for (int k = 0; k < KM; k++) {
        const int32_t e = k < m.K ? m.ell[k * NI + c] : -1;
        fk[k] = e >= 0 ? (int64_t)(e >> 1) : 0;
        sg[k] = e < 0 ? 0.0 : ((e & 1) ? 1.0 : -1.0);
    }
    double v[KM][FS];
#pragma unroll
    for (int k = 0; k < KM; k++) {
        const double2* rec = reinterpret_cast<const double2*>(Fe + FS * fk[k]);
#pragma unroll
        for (int i = 0; i < FS / 2; i++) {
            const double2 t = rec[i];
            v[k][2 * i] = t.x;
            v[k][2 * i + 1] = t.y;
        }
    }
    double acc[NE];
#pragma unroll
    for (int p = 0; p < NE; p++) {
        acc[p] = 0.0;
#pragma unroll
        for (int k = 0; k < KM; k++)
            if (sg[k] != 0.0) acc[p] += sg[k] * v[k][p];
    }
    const double iv = m.inv_vol[c];
#pragma unroll
    for (int p = 0; p < NE; p++) {
        const double r = acc[p] * iv;
        R[p * N + c] = accumulate ? R[p * N + c] + r : r;
    }
    if (!SPECIES) return;
    // species pairs (sp, sp + 1) interleaved per face: one 16-byte load per face and pair
    const double2* __restrict__ Fs2 = reinterpret_cast<const double2*>(Fs);
    for (int q = 0; 2 * q < ns; q++) {
        double a0 = 0.0, a1 = 0.0;
#pragma unroll
        for (int k = 0; k < KM; k++) {
            const double2 fv = Fs2[(int64_t)q * F + fk[k]];
            if (sg[k] != 0.0) {
                a0 += sg[k] * fv.x;
                a1 += sg[k] * fv.y;
            }
        }
        for (int h = 0; h < 2 && 2 * q + h < ns; h++) {
            const int sp = 2 * q + h;
            const double a = h ? a1 : a0;
            const int64_t comp = NE + sp;
            const double r = a * iv;
            R[comp * N + c] = accumulate ? R[comp * N + c] + r : r;
        }
    }
}

#include "fv_fused.hpp"

// ============================================================================================ mixed form
//
// The residual of a cell is a sum of face fluxes that cancel to O(1 %) of their size (and far more at low Mach),
// so FP32 fluxes would lose ~2 digits.  The mixed form therefore evaluates, per face, the PERTURBATION of the flux
// against the flux of the owning cell's own state, dF = F(U_L, U_R) - F(U_c).n: the baseline terms cancel exactly
// over a closed cell (sum_f n_f A_f = 0), and dF is small, so FP32 keeps ~7 digits of it.  The cell state is
// decoded in FP64 (pressure / temperature recovery from total energy); face-side states enter as FP32 perturbations
// dU (FP64 differences of cell states rounded to FP32, plus FP32 reconstruction increments), turned into
// perturbations of the primitive variables by exact identities; the AUSM split functions enter through factored
// differences, e.g. P+(M_L) - P+(M_c) = 1/4 (M_L - M_c)(3 - M_L^2 - M_L M_c - M_c^2).  Faces with a supersonic
// side, and the non-AUSM calculators, take the FP64 route (full flux minus FP64 baseline).  Accumulation is FP64.

template <int DIM>
struct Baseline {  // FP64 EOS recovery of a cell (perfectGas.cpp:67-113, 147-174)
    double rho, vel[DIM], eps, T, p, a, H;
};

template <int DIM>
__device__ inline void decode_baseline(const double* u, const Phys& ph, Baseline<DIM>& b) {
    b.rho = u[kRho];
    const double irho = 1.0 / b.rho;
    double ke = 0.0;
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        b.vel[d] = u[kRhoU + d] * irho;
        ke += b.vel[d] * b.vel[d];
    }
    b.eps = u[kRhoE] * irho;
    b.T = (b.eps - 0.5 * ke) * ph.icv;
    b.p = ph.R * b.rho * b.T;
    b.a = sqrt(ph.gamma * ph.R * b.T);
    b.H = b.eps + ph.R * b.T;
}

template <int DIM>
struct BaselineF {  // FP32 copies of the baseline for products with perturbations
    float rho, vel[DIM], eps, T, p, a, H;
};

template <int DIM>
struct SidePert {
    float drho, rho, du[DIM], dun, dT, dp, a, da, dM, M, dH;
};

// Perturbations of the primitive variables of a face side from its conserved perturbation dU.
template <int DIM>
__device__ inline void side_perturbation(const float* dU, const BaselineF<DIM>& b, const float* nhat, float unc, float Mc,
                                         float gR, float R, float icv, SidePert<DIM>& sp) {
    sp.drho = dU[kRho];
    sp.rho = b.rho + sp.drho;
    const float irho = 1.0f / sp.rho;
    float dun = 0.f, dke = 0.f, dd = 0.f;
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        sp.du[d] = (dU[kRhoU + d] - b.vel[d] * sp.drho) * irho;
        dun += sp.du[d] * nhat[d];
        dke += b.vel[d] * sp.du[d];
        dd += sp.du[d] * sp.du[d];
    }
    const float deps = (dU[kRhoE] - b.eps * sp.drho) * irho;
    const float de = deps - (dke + 0.5f * dd);
    sp.dT = de * icv;
    sp.dp = R * (sp.rho * sp.dT + b.T * sp.drho);
    sp.a = sqrtf(gR * (b.T + sp.dT));
    sp.da = gR * sp.dT / (sp.a + b.a);
    sp.dun = dun;
    sp.dM = (dun * b.a - unc * sp.da) / (sp.a * b.a);
    sp.M = Mc + sp.dM;
    sp.dH = deps + R * sp.dT;
}

// Perturbation of the physical flux f.n = (m, m H, m u + p n), m = rho u_n, for a side perturbation sp of the
// baseline (conserved perturbation already decoded by side_perturbation): dm = drho u_n + rho' du_n,
// d(m H) = dm H + m' dH, d(m u + p n) = dm u + m' du + dp n.
template <int DIM>
__device__ inline void phys_flux_pert(const SidePert<DIM>& sp, const BaselineF<DIM>& b, float unc, float mdot_c,
                                      const float* nhat, float* out) {
    const float dm = sp.drho * unc + sp.rho * sp.dun;
    const float mp = mdot_c + dm;
    out[kRho] = dm;
    out[kRhoE] = dm * b.H + mp * sp.dH;
#pragma unroll
    for (int d = 0; d < DIM; d++) out[kRhoU + d] = dm * b.vel[d] + mp * sp.du[d] + sp.dp * nhat[d];
}

// Mixed-form flux of face f with both baselines: dFL = F - F(U_L).n A and dFR = F - F(U_R).n A (undefined for a
// ghost right cell).  The face flux is formed as a perturbation of the left cell's state; the right baseline
// differs by the physical-flux perturbation of the cell difference U_R - U_L, also formed in FP32 from exact
// identities (or in FP64 on the FP64 route).  RUS: the Rusanov calculator, always in FP32 perturbation form (it has
// no supersonic branch), compiled without the FP64 route.
template <int DIM, bool RECON, bool VISC, bool SPECIES, bool RUS>
__device__ inline void mixed_face(const MeshView<float>& m, const Phys& ph, const double* __restrict__ U,
                                  const float* __restrict__ rec, const float* __restrict__ gS,
                                  const float* __restrict__ gY, int64_t f, float* dFL, float* dFR, float* sFL,
                                  float* sFR, int64_t s_stride) {
    constexpr int NE = DIM + 2;
    constexpr int NT = ViscTerms<DIM>::NT;
    constexpr bool USE_REC = RECON || VISC;
    const int64_t N = m.N, F = m.F, NI = m.NI;
    const int64_t L = m.fcell[f], Rc = m.fcell[F + f];
    const bool ghost = Rc >= NI;
    float nh[3], area, x[2][DIM];
    load_face_na(m.fna, f, nh, area);
    load_face_pair<DIM, float>(m.fdx, f, x);
    float nhat[DIM];
#pragma unroll
    for (int d = 0; d < DIM; d++) nhat[d] = nh[d];
    RecF<DIM, RECON, VISC> rl, rr;
    if (USE_REC) {
        load_rec_f<DIM, RECON, VISC>(rec, L, rl);
    } else {
#pragma unroll
        for (int p = 0; p < NE; p++) rl.u[p] = U[p * N + L];
    }
    if (USE_REC && !ghost) {
        load_rec_f<DIM, RECON, VISC>(rec, Rc, rr);
    } else {
#pragma unroll
        for (int p = 0; p < NE; p++) rr.u[p] = U[p * N + Rc];
    }
    Baseline<DIM> b;
    decode_baseline<DIM>(rl.u, ph, b);
    BaselineF<DIM> bf;
    bf.rho = float(b.rho);
    bf.eps = float(b.eps);
    bf.T = float(b.T);
    bf.p = float(b.p);
    bf.a = float(b.a);
    bf.H = float(b.H);
#pragma unroll
    for (int d = 0; d < DIM; d++) bf.vel[d] = float(b.vel[d]);
    const float gR = float(ph.gamma * ph.R), Rg = float(ph.R), icv = float(ph.icv);
    // conserved perturbations relative to the left cell: the cell difference and both face states
    float dcell[NE], dL[NE], dR[NE];
#pragma unroll
    for (int p = 0; p < NE; p++) {
        dcell[p] = float(rr.u[p] - rl.u[p]);
        float a = 0.f, bb = 0.f;
        if (RECON)
#pragma unroll
            for (int d = 0; d < DIM; d++) {
                a += rl.g[p * DIM + d] * x[0][d];
                if (!ghost) bb += rr.g[p * DIM + d] * x[1][d];
            }
        dL[p] = a;
        dR[p] = dcell[p] + bb;
    }
    float unc = 0.f;
#pragma unroll
    for (int d = 0; d < DIM; d++) unc += bf.vel[d] * nhat[d];
    const float Mc = unc / bf.a;
    SidePert<DIM> sL, sR;
    side_perturbation<DIM>(dL, bf, nhat, unc, Mc, gR, Rg, icv, sL);
    side_perturbation<DIM>(dR, bf, nhat, unc, Mc, gR, Rg, icv, sR);
    float dF[NE], dPhi[NE];
#pragma unroll
    for (int p = 0; p < NE; p++) {
        dF[p] = 0.f;
        dPhi[p] = 0.f;
    }
    const bool subsonic = fabsf(sL.M) <= 1.f && fabsf(sR.M) <= 1.f && fabsf(Mc) <= 1.f;
    // species advection coefficients, set by the flux branch and applied per species after the viscous part
    // (kind 0: no advection, 1: Rusanov, 2: subsonic AUSM, 3: FP64 route)
    int sp_kind = 0;
    float sp_a = 0.f, sp_b = 0.f, sp_c = 0.f, sp_d = 0.f, sp_e = 0.f, sp_scdun = 0.f;
    bool sp_upR = false;
    double sp_clo = 0.0, sp_cro = 0.0, sp_und = 0.0, sp_unr = 0.0;
    if constexpr (RUS) {
        // Rusanov (build-defined) as a perturbation of the left baseline: F - f(U_L).n A =
        // A [1/2 (df(dL) + df(dR)) - 1/2 lambda (dR - dL)], lambda = max(|u_n| + a) of the face states
        const float mdot_c = unc * bf.rho;
        float dfl[NE], dfr[NE];
        phys_flux_pert<DIM>(sL, bf, unc, mdot_c, nhat, dfl);
        phys_flux_pert<DIM>(sR, bf, unc, mdot_c, nhat, dfr);
        const float unL = unc + sL.dun, unR = unc + sR.dun;
        const float spL = fabsf(unL) + sL.a, spR = fabsf(unR) + sR.a;
        const float lam = spL < spR ? spR : spL;  // pmax order, as rusanov_flux
#pragma unroll
        for (int p = 0; p < NE; p++) dF[p] = area * (0.5f * (dfl[p] + dfr[p]) - 0.5f * lam * (dR[p] - dL[p]));
        SidePert<DIM> sC;
        if (!ghost) {
            side_perturbation<DIM>(dcell, bf, nhat, unc, Mc, gR, Rg, icv, sC);
            phys_flux_pert<DIM>(sC, bf, unc, mdot_c, nhat, dPhi);
#pragma unroll
            for (int p = 0; p < NE; p++) dPhi[p] *= area;
        }
        if (SPECIES) {
            // F_sp = A [1/2 (rhoY_L u_n,L + rhoY_R u_n,R) - 1/2 lambda (rhoY_R - rhoY_L)] relative to rhoY_c u_n,c:
            // A [rhoY_c 1/2 (du_n,L + du_n,R) + (1/2 u_n,L + 1/2 lambda) dsL + (1/2 u_n,R - 1/2 lambda) dsR]
            sp_kind = 1;
            sp_a = 0.5f * unL + 0.5f * lam;   // cl
            sp_b = 0.5f * unR - 0.5f * lam;   // cr
            sp_c = 0.5f * (sL.dun + sR.dun);  // dun2
            sp_scdun = ghost ? 0.f : sC.dun;
        }
    } else if (ph.flux == ABL_FV_FLUX_AUSM && subsonic) {
        const float dm = 0.25f * sL.dM * (sL.M + Mc + 2.f) - 0.25f * sR.dM * (sR.M + Mc - 2.f);
        const float mm = Mc + dm;
        const SidePert<DIM>& up = mm < 0.f ? sR : sL;
        const float mdot_c = unc * bf.rho;
        const float dmdot = dm * up.a * up.rho + Mc * (up.da * up.rho + bf.a * up.drho);
        const float mdot = mdot_c + dmdot;
        const float PpL = 0.25f * (sL.M + 1.f) * (sL.M + 1.f) * (2.f - sL.M);
        const float PmR = 0.25f * (sR.M - 1.f) * (sR.M - 1.f) * (2.f + sR.M);
        const float dPpL = 0.25f * sL.dM * (3.f - (sL.M * sL.M + sL.M * Mc + Mc * Mc));
        const float dPpR = 0.25f * sR.dM * (3.f - (sR.M * sR.M + sR.M * Mc + Mc * Mc));
        const float dp12 = PpL * sL.dp + PmR * sR.dp + bf.p * (dPpL - dPpR);
        dF[kRho] = dmdot;
        dF[kRhoE] = bf.H * dmdot + up.dH * mdot;
#pragma unroll
        for (int d = 0; d < DIM; d++) dF[kRhoU + d] = bf.vel[d] * dmdot + up.du[d] * mdot + dp12 * nhat[d];
#pragma unroll
        for (int p = 0; p < NE; p++) dF[p] *= area;
        if (!ghost) {
            // F(U_R).n - F(U_L).n from the cell difference: m = rho u_n, F = (m, m H, m u + p n)
            SidePert<DIM> sC;
            side_perturbation<DIM>(dcell, bf, nhat, unc, Mc, gR, Rg, icv, sC);
            phys_flux_pert<DIM>(sC, bf, unc, mdot_c, nhat, dPhi);
#pragma unroll
            for (int p = 0; p < NE; p++) dPhi[p] *= area;
        }
        if (SPECIES) {
            // F_sp = mdot Y_up A (speciesTransport.cpp:325-337) as a perturbation of the left baseline rhoY_L u_n
            SidePert<DIM> sC;
            if (!ghost) side_perturbation<DIM>(dcell, bf, nhat, unc, Mc, gR, Rg, icv, sC);
            sp_kind = 2;
            sp_upR = mm < 0.f;
            sp_a = up.drho;
            sp_b = up.rho;
            sp_c = dmdot;
            sp_d = mdot;
            sp_scdun = ghost ? 0.f : sC.dun;
        }
    } else if (ph.flux != ABL_FV_FLUX_NONE && ph.flux != ABL_FV_FLUX_OFF && ph.flux != ABL_FV_FLUX_RUSANOV) {
        // FP64 route: full face flux of the reconstructed states minus the FP64 baselines
        double uLd[NE], uRd[NE], nhd[DIM], fl[NE];
#pragma unroll
        for (int p = 0; p < NE; p++) {
            uLd[p] = rl.u[p] + double(dL[p]);
            uRd[p] = rl.u[p] + double(dR[p]);
            fl[p] = 0.0;
        }
#pragma unroll
        for (int d = 0; d < DIM; d++) nhd[d] = double(nhat[d]);
        const double ad = double(area);
        double clo = 0.0, cro = 0.0;  // species flux coefficients: F_sp = clo rhoY_L,f + cro rhoY_R,f
        {
            double mA, rlo, rro;
            int dir;
            advection_flux<double, DIM, false>(ph.flux, uLd, uRd, nhd, ad, ph, fl, mA, dir, rlo, rro);
            species_coefficients<double>(dir, ph.compact != 0, mA, rlo, rro, clo, cro);
        }
        double und = 0.0;
#pragma unroll
        for (int d = 0; d < DIM; d++) und += b.vel[d] * nhd[d];
        const double mc = b.rho * und * ad;
        double base[NE];
        base[kRho] = mc;
        base[kRhoE] = b.H * mc;
#pragma unroll
        for (int d = 0; d < DIM; d++) base[kRhoU + d] = b.vel[d] * mc + b.p * ad * nhd[d];
#pragma unroll
        for (int p = 0; p < NE; p++) dF[p] = float(fl[p] - base[p]);
        if (!ghost) {
            Baseline<DIM> br;
            decode_baseline<DIM>(rr.u, ph, br);
            double unr = 0.0;
#pragma unroll
            for (int d = 0; d < DIM; d++) unr += br.vel[d] * nhd[d];
            const double mr = br.rho * unr * ad;
            dPhi[kRho] = float(mr - base[kRho]);
            dPhi[kRhoE] = float(br.H * mr - base[kRhoE]);
#pragma unroll
            for (int d = 0; d < DIM; d++)
                dPhi[kRhoU + d] = float(br.vel[d] * mr + br.p * ad * nhd[d] - base[kRhoU + d]);
        }
        if (SPECIES) {
            double unr = 0.0;
            if (!ghost) {
                Baseline<DIM> br;
                decode_baseline<DIM>(rr.u, ph, br);
#pragma unroll
                for (int d = 0; d < DIM; d++) unr += br.vel[d] * nhd[d];
            }
            sp_kind = 3;
            sp_clo = clo;
            sp_cro = cro;
            sp_und = und;
            sp_unr = unr;
            sp_e = area;
        }
    }
    if (VISC) {
        // face-averaged viscous terms and face velocity (cell values) in FP32; a full flux, shared by both sides
        float tq[NT], vel[DIM], normal[DIM];
        const float irn = 1.0f / (bf.rho + dcell[kRho]);
        if (ghost) {
            float w[2][DIM], gA[DIM + 1][DIM];
            load_face_pair<DIM, float>(m.fwa, f, w);
            float dvel[DIM], dke = 0.f, dd = 0.f;
#pragma unroll
            for (int d = 0; d < DIM; d++) {
                dvel[d] = (dcell[kRhoU + d] - bf.vel[d] * dcell[kRho]) * irn;
                dke += bf.vel[d] * dvel[d];
                dd += dvel[d] * dvel[d];
            }
            const float dT = ((dcell[kRhoE] - bf.eps * dcell[kRho]) * irn - (dke + 0.5f * dd)) * icv;
#pragma unroll
            for (int d = 0; d < DIM; d++) {
                gA[0][d] = -w[1][d] * dT;
#pragma unroll
                for (int q = 0; q < DIM; q++) gA[1 + q][d] = -w[1][d] * dvel[q];
            }
            stress_heat<float, DIM>(float(ph.mu), float(ph.k), gA, tq);
        } else {
#pragma unroll
            for (int q = 0; q < NT; q++) tq[q] = rr.tq[q];
        }
#pragma unroll
        for (int q = 0; q < NT; q++) tq[q] = 0.5f * rl.tq[q] + 0.5f * tq[q];
#pragma unroll
        for (int d = 0; d < DIM; d++) {
            const float dv = (dcell[kRhoU + d] - bf.vel[d] * dcell[kRho]) * irn;
            vel[d] = bf.vel[d] + 0.5f * dv;
            normal[d] = area * nhat[d];
        }
        viscous_flux_add<float, DIM>(normal, vel, tq, dF);
    }
    if (SPECIES) {
        // per species: the advection perturbation of the flux branch's coefficients, plus DiffusionSpeciesFlux
        // (speciesTransport.cpp:196-228): -D rho_face n.(1/2 (gradY_L + gradY_R)) A, a full flux (no baseline)
        // shared by both sides.  Both sides' gradient words come in FP32 pairs, each read once.
        const bool sdiff = VISC && ph.sdiff;
        const float cd = sdiff ? -float(ph.diff) * (bf.rho + 0.5f * dcell[kRho]) * 0.5f * area : 0.f;
        float w[2][DIM];
        if (sdiff && ghost) load_face_pair<DIM, float>(m.fwa, f, w);
        const float2* __restrict__ gS2 = reinterpret_cast<const float2*>(gS);
        for (int sp = 0; sp < ph.ns; sp++) {
            const int64_t comp = NE + sp;
            const double yl = U[comp * N + L], yr = U[comp * N + Rc];
            float wl[2 * DIM], wr[2 * DIM];
#pragma unroll
            for (int j = 0; j < DIM; j++) {
                const bool need = (RECON && 2 * j < DIM) || (sdiff && 2 * j + 1 >= DIM);
                const float2 a = need ? gS2[((int64_t)sp * DIM + j) * N + L] : make_float2(0.f, 0.f);
                const float2 c2 = (need && !ghost) ? gS2[((int64_t)sp * DIM + j) * N + Rc] : make_float2(0.f, 0.f);
                wl[2 * j] = a.x;
                wl[2 * j + 1] = a.y;
                wr[2 * j] = c2.x;
                wr[2 * j + 1] = c2.y;
            }
            const float dcs = float(yr - yl);
            float dsl = 0.f, dsr = dcs;
            if (RECON)
#pragma unroll
                for (int d = 0; d < DIM; d++) {
                    dsl += wl[d] * x[0][d];
                    if (!ghost) dsr += wr[d] * x[1][d];
                }
            float fl = 0.f, fr = 0.f;
            if (sp_kind == 1 || sp_kind == 2) {
                float adv;
                if (sp_kind == 1) {
                    adv = area * (float(yl) * sp_c + sp_a * dsl + sp_b * dsr);
                } else {
                    const float YL = float(yl / rl.u[kRho]);
                    const float dsu = sp_upR ? dsr : dsl;
                    const float dYu = (dsu - YL * sp_a) / sp_b;
                    adv = (sp_c * YL + sp_d * dYu) * area;
                }
                float shift = 0.f;
                if (!ghost) shift = (dcs * unc + (float(yl) + dcs) * sp_scdun) * area;
                fl = adv;
                fr = adv - shift;
            } else if (sp_kind == 3) {
                const double ad = double(sp_e);
                const double fsp = sp_clo * (yl + double(dsl)) + sp_cro * (yl + double(dsr)) - yl * sp_und * ad;
                fl = float(fsp);
                fr = ghost ? 0.f : float(fsp - (yr * sp_unr - yl * sp_und) * ad);
            }
            if (sdiff) {
                float dot = 0.f;
                float gyg = 0.f;
                if (ghost) {
                    const double ylm = yl / rl.u[kRho], yg = yr / rr.u[kRho];
                    gyg = float(ylm - yg);
                }
#pragma unroll
                for (int d = 0; d < DIM; d++) {
                    const float gl = wl[DIM + d];
                    const float gr = ghost ? w[1][d] * gyg : wr[DIM + d];
                    dot += nhat[d] * (gl + gr);
                }
                fl += cd * dot;
                fr += cd * dot;
            }
            sFL[sp * s_stride] = fl;
            sFR[sp * s_stride] = fr;
        }
    }
#pragma unroll
    for (int p = 0; p < NE; p++) {
        dFL[p] = dF[p];
        dFR[p] = dF[p] - dPhi[p];
    }
}

// Mixed form, cell-centric (strategy "recompute"): each cell evaluates its faces and keeps its own side's
// perturbation; accumulation in FP64.  Euler components only (species need the tile strategy).
template <int DIM, bool RECON, bool VISC, bool RUS>
__global__ __launch_bounds__(kBlock) void k_flux_mixed(MeshView<float> m, Phys ph, const double* __restrict__ U,
                                                       const float* __restrict__ rec, double* __restrict__ R,
                                                       int accumulate) {
    constexpr int NE = DIM + 2;
    const int64_t c = xcd_cell_index();
    if (c >= m.NI) return;
    const int64_t N = m.N, NI = m.NI;
    double acc[NE];
#pragma unroll
    for (int p = 0; p < NE; p++) acc[p] = 0.0;
    for (int k = 0; k < m.K; k++) {
        const int32_t e = m.ell[k * NI + c];
        if (e < 0) break;
        float dFL[NE], dFR[NE];
        mixed_face<DIM, RECON, VISC, false, RUS>(m, ph, U, rec, nullptr, nullptr, (int64_t)(e >> 1), dFL, dFR, nullptr,
                                            nullptr, 0);
        if (e & 1) {
#pragma unroll
            for (int p = 0; p < NE; p++) acc[p] += double(dFR[p]);
        } else {
#pragma unroll
            for (int p = 0; p < NE; p++) acc[p] -= double(dFL[p]);
        }
    }
    const double iv = m.inv_vol[c];
#pragma unroll
    for (int p = 0; p < NE; p++) {
        const double r = acc[p] * iv;
        R[p * N + c] = accumulate ? R[p * N + c] + r : r;
    }
}

// Mixed form, tile strategy: each face of a tile evaluated once into LDS (both baselines, FP32: rows [0, NE + ns)
// left-baseline perturbations, rows [NE + ns, 2 (NE + ns)) right-baseline ones), gathered per cell in FP64 in
// face-number order.
template <int DIM, bool RECON, bool VISC, bool SPECIES, int KM, bool RUS>
__global__ __launch_bounds__(kTileMax) void k_tile_flux_mixed(MeshView<float> m, Phys ph, const double* __restrict__ U,
                                                              const float* __restrict__ rec,
                                                              const float* __restrict__ gS,
                                                              const float* __restrict__ gY,
                                                              const int32_t* __restrict__ tstart,
                                                              const int32_t* __restrict__ tfaces,
                                                              const int32_t* __restrict__ tslot, int nfm,
                                                              double* __restrict__ R, int accumulate) {
    constexpr int NE = DIM + 2;
    extern __shared__ double lds_d[];
    float* lf = reinterpret_cast<float*>(lds_d);
    const int nrow = NE + (SPECIES ? ph.ns : 0);
    const int64_t t = xcd_block_index();
    const int32_t f0 = tstart[t], nf = tstart[t + 1] - f0;
    for (int i = threadIdx.x; i < nf; i += blockDim.x) {
        float dFL[NE], dFR[NE];
        mixed_face<DIM, RECON, VISC, SPECIES, RUS>(m, ph, U, rec, gS, gY, (int64_t)tfaces[f0 + i], dFL, dFR,
                                              lf + NE * nfm + i, lf + (nrow + NE) * nfm + i, nfm);
#pragma unroll
        for (int p = 0; p < NE; p++) {
            lf[p * nfm + i] = dFL[p];
            lf[(nrow + p) * nfm + i] = dFR[p];
        }
    }
    __syncthreads();
    const int64_t c = t * blockDim.x + threadIdx.x;
    if (c >= m.NI) return;
    const int64_t N = m.N, NI = m.NI;
    int off[KM];
    double sg[KM];
#pragma unroll
    for (int k = 0; k < KM; k++) {
        const int32_t e = tslot[k * NI + c];
        off[k] = e >= 0 ? ((e & 1) ? nrow * nfm : 0) + (e >> 1) : 0;
        sg[k] = e < 0 ? 0.0 : ((e & 1) ? 1.0 : -1.0);
    }
    const double iv = m.inv_vol[c];
    for (int p = 0; p < nrow; p++) {
        double acc = 0.0;
#pragma unroll
        for (int k = 0; k < KM; k++)
            if (sg[k] != 0.0) acc += sg[k] * double(lf[p * nfm + off[k]]);
        const double r = acc * iv;
        R[p * N + c] = accumulate ? R[p * N + c] + r : r;
    }
}

// Mixed form, face-flux strategy, pass 1: one lane per face stores both sides' FP32 perturbations as rows
// [2 nrow][F] (left-baseline rows, then right-baseline rows; nrow = n_euler + n_species): coalesced stores.  Each
// species perturbation is formed once (advection and diffusion together) and stored straight into its rows.
template <int DIM, bool RECON, bool VISC, bool SPECIES, bool RUS>
__global__ __launch_bounds__(kBlock) void k_face_flux_mixed(MeshView<float> m, Phys ph, const double* __restrict__ U,
                                                            const float* __restrict__ rec, const float* __restrict__ gS,
                                                            const float* __restrict__ gY, float* __restrict__ Fm,
                                                            int nrow) {
    constexpr int NE = DIM + 2;
    const int64_t f = xcd_cell_index();
    if (f >= m.F) return;
    const int64_t F = m.F;
    float dFL[NE], dFR[NE];
    mixed_face<DIM, RECON, VISC, SPECIES, RUS>(m, ph, U, rec, gS, gY, f, dFL, dFR, Fm + NE * F + f,
                                               Fm + (nrow + NE) * F + f, F);
#pragma unroll
    for (int p = 0; p < NE; p++) {
        Fm[p * F + f] = dFL[p];
        Fm[(nrow + p) * F + f] = dFR[p];
    }
}

// Mixed form, face-flux strategy, pass 2: each interior cell gathers its own side's perturbations in face-number
// order, in FP64; the same sums in the same order as the tile gather, so both strategies give the same residual.
template <int DIM, int KM>
__global__ __launch_bounds__(kBlock) void k_gather_mixed(MeshView<float> m, int nrow, const float* __restrict__ Fm,
                                                         double* __restrict__ R, int accumulate) {
    const int64_t c = xcd_cell_index();
    if (c >= m.NI) return;
    const int64_t N = m.N, NI = m.NI;
    int64_t off[KM];
    double sg[KM];
#pragma unroll
    for (int k = 0; k < KM; k++) {
        const int32_t e = k < m.K ? m.ell[k * NI + c] : -1;
        off[k] = e >= 0 ? (int64_t)(e >> 1) + ((e & 1) ? (int64_t)nrow * m.F : 0) : 0;
        sg[k] = e < 0 ? 0.0 : ((e & 1) ? 1.0 : -1.0);
    }
    const double iv = m.inv_vol[c];
    for (int p = 0; p < nrow; p++) {
        double acc = 0.0;
#pragma unroll
        for (int k = 0; k < KM; k++)
            if (sg[k] != 0.0) acc += sg[k] * double(Fm[p * m.F + off[k]]);
        const double r = acc * iv;
        R[p * N + c] = accumulate ? R[p * N + c] + r : r;
    }
}

// ============================================================================================ adjoint
//
// Vector-Jacobian product of the first-order (no reconstruction) inviscid RHS: Ubar = Rbar^T dR/dU and
// theta_bar = Rbar^T dR/dtheta for inflow ghosts, theta = (rho_in, |u_in|, T_in).  R_c = (1/V_c) sum_f sign F_f, so
// every face contributes dF/dU_side^T g_f with g_f = Rbar_R / V_R - Rbar_L / V_L.  One lane per interior cell
// recomputes the forward flux of each of its faces and keeps its own side's adjoint (deterministic gather).

// Adjoint seeds of face f: g[p] for euler components and the species sums QL = sum g_sp rhoY_L,
// QR = sum g_sp rhoY_R, QdL = sum g_sp (rhoY_R - rhoY_L).
template <int DIM, int UNR = 2>
__device__ inline void face_seed(const MeshView<double>& m, const Phys& ph, const double* __restrict__ U,
                                 const double* __restrict__ Rb, int64_t L, int64_t Rc, bool ghost, double* fb,
                                 double& QL, double& QR, double& QdL, double* gout = nullptr, int64_t gstride = 0) {
    constexpr int NE = DIM + 2;
    const int64_t N = m.N;
    const double ivL = m.inv_vol[L], ivR = ghost ? 0.0 : m.inv_vol[Rc];
#pragma unroll
    for (int p = 0; p < NE; p++) fb[p] = (ghost ? 0.0 : Rb[p * N + Rc] * ivR) - Rb[p * N + L] * ivL;
    QL = 0.0;
    QR = 0.0;
    QdL = 0.0;
#pragma unroll UNR
    for (int sp = 0; sp < ph.ns; sp++) {  // UNR species' loads in flight
        const int64_t comp = NE + sp;
        const double g = (ghost ? 0.0 : Rb[comp * N + Rc] * ivR) - Rb[comp * N + L] * ivL;
        if (gout) gout[sp * gstride] = g;
        const double yl = U[comp * N + L], yr = U[comp * N + Rc];
        QL += g * yl;
        QR += g * yr;
        QdL += g * (yr - yl);
    }
}

template <int DIM, bool RUS>
__device__ inline void face_vjp(const double* uL, const double* uR, double QL, double QR, double QdL, const double* nh,
                                double area, const Phys& ph, const double* fb, double* ubL, double* ubR, double& cL,
                                double& cR) {
    if (RUS)
        rusanov_flux_vjp<DIM>(uL, uR, QL, QR, QdL, nh, area, ph, fb, ubL, ubR, cL, cR);
    else
        ausm_flux_vjp<DIM>(uL, uR, QL, QR, nh, area, ph, fb, ubL, ubR, cL, cR);
}

template <int DIM, bool RUS, int KM>
__global__ __launch_bounds__(kBlock) void k_vjp(MeshView<double> m, Phys ph, const double* __restrict__ U,
                                                const double* __restrict__ Rb, const int32_t* __restrict__ ghost_slot,
                                                const int32_t* __restrict__ slot_kind, double* __restrict__ Ub) {
    constexpr int NE = DIM + 2;
    const int64_t c = xcd_cell_index();
    if (c >= m.NI) return;
    const int64_t N = m.N, F = m.F, NI = m.NI;
    double ub[NE], coef[KM], cghost[KM];  // cghost: species multiplier of a caller-filled ghost across slot k
#pragma unroll
    for (int p = 0; p < NE; p++) ub[p] = 0.0;
#pragma unroll
    for (int k = 0; k < KM; k++) {
        const Slot sl = face_slot(m, k, c);
        const int64_t f = sl.f;
        const int64_t L = m.fcell[f], Rc = m.fcell[F + f];
        const bool ghost = Rc >= NI;
        cghost[k] = 0.0;
        double nh[3], area, uL[NE], uR[NE], fb[NE], QL, QR, QdL;
        load_face_na(m.fna, f, nh, area);
#pragma unroll
        for (int p = 0; p < NE; p++) {
            uL[p] = U[p * N + L];
            uR[p] = U[p * N + Rc];
        }
        face_seed<DIM>(m, ph, U, Rb, L, Rc, ghost, fb, QL, QR, QdL);
        double ubL[NE], ubR[NE], cL, cR;
#pragma unroll
        for (int p = 0; p < NE; p++) {
            ubL[p] = 0.0;
            ubR[p] = 0.0;
        }
        face_vjp<DIM, RUS>(uL, uR, QL, QR, QdL, nh, area, ph, fb, ubL, ubR, cL, cR);
        coef[k] = 0.0;
        if (!sl.valid) continue;
        double cf = sl.s == 0 ? cL : cR;
#pragma unroll
        for (int p = 0; p < NE; p++) ub[p] += sl.s == 0 ? ubL[p] : ubR[p];
        if (ghost) {
            const int gk = slot_kind[ghost_slot[Rc - NI]];
            if (gk == ABL_FV_BC_OUTFLOW) {  // the ghost copies this cell's state: its adjoint flows back here
#pragma unroll
                for (int p = 0; p < NE; p++) ub[p] += ubR[p];
                cf += cR;
            } else if (gk == kBcInflowAtFace) {  // ghost = 2 state - this cell
#pragma unroll
                for (int p = 0; p < NE; p++) ub[p] -= ubR[p];
                cf -= cR;
            } else if (gk == ABL_FV_BC_GHOST) {  // caller-filled ghost: its own adjoint row (one face per ghost)
#pragma unroll
                for (int p = 0; p < NE; p++) Ub[p * N + Rc] = ubR[p];
                cghost[k] = cR;
            }
        }
        coef[k] = cf;
    }
#pragma unroll
    for (int p = 0; p < NE; p++) Ub[p * N + c] = ub[p];
    if (ph.ns == 0) return;
    // species: the faces' cells and inverse volumes are loaded once for all species
    int64_t sL[KM], sR[KM];
    double ivL[KM], ivR[KM];
#pragma unroll
    for (int k = 0; k < KM; k++) {
        const Slot sl = face_slot(m, k, c);
        sL[k] = m.fcell[sl.f];
        sR[k] = m.fcell[F + sl.f];
        ivL[k] = m.inv_vol[sL[k]];
        ivR[k] = sR[k] >= NI ? 0.0 : m.inv_vol[sR[k]];
        if (!sl.valid) coef[k] = 0.0;
    }
#pragma unroll 2
    for (int sp = 0; sp < ph.ns; sp++) {
        const int64_t comp = NE + sp;
        double a = 0.0;
#pragma unroll
        for (int k = 0; k < KM; k++) {
            const bool ghost = sR[k] >= NI;
            const double g = (ghost ? 0.0 : Rb[comp * N + sR[k]] * ivR[k]) - Rb[comp * N + sL[k]] * ivL[k];
            a += coef[k] * g;
            if (cghost[k] != 0.0) Ub[comp * N + sR[k]] = cghost[k] * g;
        }
        Ub[comp * N + c] = a;
    }
}

// First-order VJP, face-centric (abl_fv_rhs_vjp; k_vjp above stays the allocation-free form): one lane per face
// reverses the face's flux once for both sides and stores the side adjoints, the species multipliers and the
// species seeds as rows [2 NE + 2 + ns][F] (ubL, ubR, cL, cR, g per species): coalesced stores.  k_vjp_gather then
// sums them per cell in slot order exactly as k_vjp sums its own evaluations, so both forms give the same Ubar bit
// for bit; each face is reversed once instead of once per side, and the gather reads each face's species seeds
// instead of both cells' Rbar rows.
template <int DIM, bool RUS>
__global__ __launch_bounds__(kBlock) void k_vjp_faces(MeshView<double> m, Phys ph, const double* __restrict__ U,
                                                      const double* __restrict__ Rb, double* __restrict__ fr) {
    constexpr int NE = DIM + 2;
    const int64_t f = xcd_cell_index();
    if (f >= m.F) return;
    const int64_t N = m.N, F = m.F, NI = m.NI;
    const int64_t L = m.fcell[f], Rc = m.fcell[F + f];
    const bool ghost = Rc >= NI;
    double nh[3], area, uL[NE], uR[NE], fb[NE], QL, QR, QdL;
    load_face_na(m.fna, f, nh, area);
#pragma unroll
    for (int p = 0; p < NE; p++) {
        uL[p] = U[p * N + L];
        uR[p] = U[p * N + Rc];
    }
    face_seed<DIM, 5>(m, ph, U, Rb, L, Rc, ghost, fb, QL, QR, QdL, fr + (2 * NE + 2) * F + f, F);
    double ubL[NE], ubR[NE], cL, cR;
#pragma unroll
    for (int p = 0; p < NE; p++) {
        ubL[p] = 0.0;
        ubR[p] = 0.0;
    }
    face_vjp<DIM, RUS>(uL, uR, QL, QR, QdL, nh, area, ph, fb, ubL, ubR, cL, cR);
#pragma unroll
    for (int p = 0; p < NE; p++) {
        fr[p * F + f] = ubL[p];
        fr[(NE + p) * F + f] = ubR[p];
    }
    fr[2 * NE * F + f] = cL;
    fr[(2 * NE + 1) * F + f] = cR;
}

// Per-cell sums of the face rows of k_vjp_faces: its own side of each face in slot order, the ghost side folded back
// for outflow (copy) and enforce-at-face inflow (2 state - cell) ghosts, a caller-filled ghost's own adjoint row;
// then each species from the faces' multipliers and seeds (the species flux is the multiplier times the species
// value).  Padding slots add 0 x (face 0's seed), as k_vjp does.
template <int DIM, int KM>
__global__ __launch_bounds__(kBlock) void k_vjp_gather(MeshView<double> m, Phys ph, const double* __restrict__ Rb,
                                                       const double* __restrict__ fr,
                                                       const int32_t* __restrict__ ghost_slot,
                                                       const int32_t* __restrict__ slot_kind, double* __restrict__ Ub) {
    constexpr int NE = DIM + 2;
    const int64_t c = xcd_cell_index();
    if (c >= m.NI) return;
    const int64_t N = m.N, F = m.F, NI = m.NI;
    double ub[NE], coef[KM], cghost[KM];
    int64_t sf[KM], sR[KM];
#pragma unroll
    for (int p = 0; p < NE; p++) ub[p] = 0.0;
#pragma unroll
    for (int k = 0; k < KM; k++) {
        const Slot sl = face_slot(m, k, c);
        const int64_t f = sl.f;
        sf[k] = f;
        sR[k] = m.fcell[F + f];
        cghost[k] = 0.0;
        coef[k] = 0.0;
        if (!sl.valid) continue;
        const int64_t Rc = sR[k];
        double cf = fr[(2 * NE + sl.s) * F + f];
#pragma unroll
        for (int p = 0; p < NE; p++) ub[p] += fr[(sl.s * NE + p) * F + f];
        if (Rc >= NI) {
            const int gk = slot_kind[ghost_slot[Rc - NI]];
            if (gk == ABL_FV_BC_OUTFLOW) {
#pragma unroll
                for (int p = 0; p < NE; p++) ub[p] += fr[(NE + p) * F + f];
                cf += fr[(2 * NE + 1) * F + f];
            } else if (gk == kBcInflowAtFace) {
#pragma unroll
                for (int p = 0; p < NE; p++) ub[p] -= fr[(NE + p) * F + f];
                cf -= fr[(2 * NE + 1) * F + f];
            } else if (gk == ABL_FV_BC_GHOST) {
#pragma unroll
                for (int p = 0; p < NE; p++) Ub[p * N + Rc] = fr[(NE + p) * F + f];
                cghost[k] = fr[(2 * NE + 1) * F + f];
            }
        }
        coef[k] = cf;
    }
#pragma unroll
    for (int p = 0; p < NE; p++) Ub[p * N + c] = ub[p];
    if (ph.ns == 0) return;
    const double* __restrict__ gs = fr + (2 * NE + 2) * F;
#pragma unroll 2
    for (int sp = 0; sp < ph.ns; sp++) {
        const int64_t comp = NE + sp;
        double a = 0.0;
#pragma unroll
        for (int k = 0; k < KM; k++) {
            const double g = gs[(int64_t)sp * F + sf[k]];
            a += coef[k] * g;
            if (cghost[k] != 0.0) Ub[comp * N + sR[k]] = cghost[k] * g;
        }
        Ub[comp * N + c] = a;
    }
}

// theta_bar contributions of each inflow boundary face: dU_ghost/dtheta^T (dF/dU_ghost)^T g_f.
// slot_param[slot] = {rho_in, |u_in|, T_in, dir_x, dir_y, dir_z}; slot_Y[slot][ns] = inflow mass fractions.
template <int DIM, bool RUS>
__global__ __launch_bounds__(kBlock) void k_vjp_theta(MeshView<double> m, Phys ph, const double* __restrict__ U,
                                                      const double* __restrict__ Rb, int64_t nb,
                                                      const int32_t* __restrict__ bface, const int32_t* __restrict__ bslot,
                                                      const int32_t* __restrict__ slot_kind,
                                                      const double* __restrict__ slot_param,
                                                      const double* __restrict__ slot_Y, double* __restrict__ contrib) {
    constexpr int NE = DIM + 2;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nb) return;
    const int64_t f = bface[i];
    const int slot = bslot[i];
    double t0 = 0.0, t1 = 0.0, t2 = 0.0;
    const int kind = slot_kind[slot];
    if (kind == ABL_FV_BC_INFLOW || kind == kBcInflowAtFace) {
        const int64_t N = m.N, F = m.F;
        const int64_t L = m.fcell[f], G = m.fcell[F + f];
        double nh[3], area, uL[NE], uG[NE], fb[NE], QL, QR, QdL;
        load_face_na(m.fna, f, nh, area);
#pragma unroll
        for (int p = 0; p < NE; p++) {
            uL[p] = U[p * N + L];
            uG[p] = U[p * N + G];
        }
        face_seed<DIM>(m, ph, U, Rb, L, G, true, fb, QL, QR, QdL);
        double ubL[NE], ubG[NE], cL, cR;
#pragma unroll
        for (int p = 0; p < NE; p++) {
            ubL[p] = 0.0;
            ubG[p] = 0.0;
        }
        face_vjp<DIM, RUS>(uL, uG, QL, QR, QdL, nh, area, ph, fb, ubL, ubG, cL, cR);
        const double* sp_ = slot_param + 6 * slot;
        const double rho = sp_[0], um = sp_[1], T = sp_[2];
        // U_ghost = [rho, rho (cv T + u^2/2), rho u dir, rho Y]
        t0 = ubG[kRho] + ubG[kRhoE] * (ph.cv * T + 0.5 * um * um);
        t1 = ubG[kRhoE] * rho * um;
        t2 = ubG[kRhoE] * rho * ph.cv;
#pragma unroll
        for (int d = 0; d < DIM; d++) {
            t0 += ubG[kRhoU + d] * um * sp_[3 + d];
            t1 += ubG[kRhoU + d] * rho * sp_[3 + d];
        }
        for (int sp = 0; sp < ph.ns; sp++) {
            const int64_t comp = NE + sp;
            const double g = -Rb[comp * N + L] * m.inv_vol[L];
            t0 += cR * g * slot_Y[slot * ph.ns + sp];
        }
        if (kind == kBcInflowAtFace) {  // ghost = 2 state(theta) - interior
            t0 *= 2.0;
            t1 *= 2.0;
            t2 *= 2.0;
        }
    }
    contrib[3 * i + 0] = t0;
    contrib[3 * i + 1] = t1;
    contrib[3 * i + 2] = t2;
}

// ============================================================================================ second-order adjoint
//
// VJP of the FP64 RHS with least-squares reconstruction (Barth-Jespersen or no limiter), the viscous flux and
// species advection / diffusion, in three passes after the forward gradient pass (cell records, species gradient
// arrays):
//   1. k_adj_face_rev (lane per face): recomputes the face flux and reverses it once for both sides into a face
//      record (face-state, viscous-term and species factors); the adjoint of a ghost state goes to a per-ghost array.
//   2. k_adj_cells_e, then k_adj_cells_s (lane per interior cell, euler field then species): each sums its faces'
//      records in slot order into the cell's adjoints (state, limited gradients through the face state U + G dx,
//      viscous terms, species aux gradients), reverses the limiter, the least-squares gradients and the aux
//      (temperature / velocity / mass fraction) gradients of the cell, and writes its own adjoint and, per face slot,
//      the contribution for the neighbour across the face.
//   3. k_adj_gather (lane per interior cell): own adjoint + the contributions its neighbours wrote for it (mirror
//      slot of each face) + the adjoint of its zero-gradient ghosts.
// Every sum runs in a fixed order (deterministic, no atomics).  maxLimGrad's cancel rule (a gradient above maxLimGrad,
// default 1e10, in a fully limited component) is taken as inactive.

struct AdjBufs {
    double* Ughost;         // [nc][NG] adjoint of ghost states
    double* Uself;          // [nc][NI]
    double* contrib;        // [KM][nc][NI]
    const int32_t* mirror;  // [KM][NI] slot of the same face in the neighbour's list, -1 when none / ghost
    int km;                 // slot width of contrib / mirror
    double* W;              // face adjoint records (AdjFaceRec pairs, 64-face blocks)
    int nwp;                // word pairs per face record
    double* GL;             // [ns][NG] densityYi adjoint of a boundary face's interior cell from the ghost's gradient
};

// Face adjoint record of the face-centric reverse pass, read back per cell by k_adj_cells_e / _s.  Words per face:
//   [0, NE)             adjoint of the left face state (the flux calculator's and AdvectionFlux's reverse)
//   [NE, 2 NE)          right face state; on a boundary face (ghost right cell, whose adjoint goes to Ughost) the
//                       left cell's extra aux adjoint [NE, NE + DIM + 1) and density adjoint (NE + DIM + 1) from the
//                       ghost's one-neighbour viscous gradient
//   [2 NE, 2 NE + NT)   half the adjoint of the face-averaged viscous terms (each side's share)
//   + DIM               half the adjoint of the face velocity
//   + 4                 species coefficients cL, cR (F_sp = cL rhoY_L,f + cR rhoY_R,f), the diffusion coefficient
//                       cd = -D rho_face A / 2 and the face density adjoint summed over species
//   + ns                per species g_sp = Rbar_sp(R) / V_R - Rbar_sp(L) / V_L
// Stored as word pairs in blocks of 64 faces (pair q of face f at double2 index (f / 64) 64 NWP + 64 q + f % 64):
// each store of a wave (64 consecutive faces) is one coalesced 1 KB access.
template <int DIM>
struct AdjFaceRec {
    static constexpr int NE = DIM + 2;
    static constexpr int NT = ViscTerms<DIM>::NT;
    static constexpr int UL = 0, UR = NE, TQ = 2 * NE, VL = TQ + NT, CL = VL + DIM, CR = CL + 1, CD = CL + 2,
                         RH = CL + 3, GS = (CL + 5) & ~1;  // species factors start on a pair boundary
    __host__ __device__ static int words(int ns) { return GS + ns; }
    __host__ __device__ static int pairs(int ns) { return (words(ns) + 1) / 2; }
};
__device__ inline int64_t adjw_pair(int64_t f, int nwp) { return (f >> 6) * 64 * (int64_t)nwp + (f & 63); }

// Reverse of one face flux (advection, viscous, species advection and diffusion), both sides at once.  One lane per
// face (interior and boundary).  The face's ghost-cell adjoints go straight to Ughost (a ghost has one face).
template <int DIM, bool RUS>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(3))) void k_adj_face_rev(MeshView<double> m, Phys ph, const double* __restrict__ U,
                                                         const double* __restrict__ Rb, const double* __restrict__ rec,
                                                         const double* __restrict__ gS, const double* __restrict__ gY,
                                                         AdjBufs b) {
    constexpr int NE = DIM + 2;
    constexpr int NT = ViscTerms<DIM>::NT;
    using CR = CellRec<DIM>;
    using AW = AdjFaceRec<DIM>;
    (void)gY;  // the Yi aux gradients live in gS's word pairs
    const int64_t f = xcd_cell_index();
    if (f >= m.F) return;
    const int64_t N = m.N, F = m.F, NI = m.NI, NG = N - NI;
    const int ns = ph.ns;
    const bool recon = ph.recon != 0, visc = ph.visc != 0, sdiff = ph.sdiff != 0;
    const bool useRec = recon || visc;
    const int64_t L = m.fcell[f], Rc = m.fcell[F + f];
    const bool ghost = Rc >= NI;
    double nh[3], area, x[2][DIM];
    load_face_na(m.fna, f, nh, area);
    load_face_pair<DIM>(m.fdx, f, x);
    const double ivL = m.inv_vol[L], ivR = ghost ? 0.0 : m.inv_vol[Rc];
    double g[NE];
#pragma unroll
    for (int p = 0; p < NE; p++) g[p] = (ghost ? 0.0 : Rb[p * N + Rc] * ivR) - Rb[p * N + L] * ivL;
    // both sides: cell state, face state and viscous terms
    double uL[NE], uR[NE], fL[NE], fR[NE], tqL[NT], tqR[NT];
#pragma unroll
    for (int sd = 0; sd < 2; sd++) {
        const int64_t cc = sd == 0 ? L : Rc;
        double* u = sd == 0 ? uL : uR;
        double* fu = sd == 0 ? fL : fR;
        double* tq = sd == 0 ? tqL : tqR;
        if (useRec && !(sd == 1 && ghost)) {
            double r[CR::RS];
            const double2* __restrict__ rb = reinterpret_cast<const double2*>(rec) + rec_pair<DIM>(cc);
#pragma unroll
            for (int q = 0; q < CR::NP; q++) {
                const double2 v = rb[CR::BLOCK * q];
                r[2 * q] = v.x;
                r[2 * q + 1] = v.y;
            }
#pragma unroll
            for (int p = 0; p < NE; p++) {
                u[p] = r[p];
                double a = 0.0;
                if (recon)
#pragma unroll
                    for (int d = 0; d < DIM; d++) a += r[CR::G0 + p * DIM + d] * x[sd][d];
                fu[p] = u[p] + a;
            }
#pragma unroll
            for (int q = 0; q < NT; q++) tq[q] = r[CR::T0 + q];
        } else {
#pragma unroll
            for (int p = 0; p < NE; p++) {
                u[p] = U[p * N + cc];
                fu[p] = u[p];
            }
#pragma unroll
            for (int q = 0; q < NT; q++) tq[q] = 0.0;
        }
    }
    // species sums of the face values; each species' g_sp = Rbar_sp(R) / V_R - Rbar_sp(L) / V_L goes straight into
    // the face record (read back below instead of the scattered Rbar rows)
    double2* __restrict__ W = reinterpret_cast<double2*>(b.W) + adjw_pair(f, b.nwp);
    constexpr int NFIX = AW::GS / 2;  // pairs holding only fixed words
    // On interior faces the species diffusion flux's density adjoint (the face density 1/2 (rho_L + rho_R) times
    // n.(grad Y_L + grad Y_R)) is summed here too, while both sides' gradient words are in registers; boundary faces
    // take it in the ghost pass below.
    double QL = 0.0, QR = 0.0, QdL = 0.0, rhob = 0.0;
    const double2* __restrict__ gS2 = reinterpret_cast<const double2*>(gS);
    for (int q = 0; 2 * q < ns; q++) {
        double gp0 = 0.0, gp1 = 0.0;
#pragma unroll 1
        for (int h = 0; h < 2; h++) {
            const int sp = 2 * q + h;
            if (sp >= ns) break;
            const int64_t comp = NE + sp;
            const double gsp = (ghost ? 0.0 : Rb[comp * N + Rc] * ivR) - Rb[comp * N + L] * ivL;
            if (h) gp1 = gsp;
            else gp0 = gsp;
            double yl = U[comp * N + L], yr = U[comp * N + Rc], dot = 0.0;
            // each side's gradient words in pairs: [0, DIM) limited densityYi gradient, [DIM, 2 DIM) Yi gradient
#pragma unroll
            for (int j = 0; j < DIM; j++) {
                const bool need = (recon && 2 * j < DIM) || (sdiff && 2 * j + 1 >= DIM);
                if (!need) continue;
                const double2 a = gS2[((int64_t)sp * DIM + j) * N + L];
                const double2 c2 = ghost ? make_double2(0.0, 0.0) : gS2[((int64_t)sp * DIM + j) * N + Rc];
#pragma unroll
                for (int h2 = 0; h2 < 2; h2++) {
                    const int w = 2 * j + h2;
                    const double wl = h2 ? a.y : a.x, wr = h2 ? c2.y : c2.x;
                    if (w < DIM) {
                        if (recon) {
                            yl += wl * x[0][w];
                            yr += wr * x[1][w];
                        }
                    } else {
                        dot += nh[w - DIM] * (wl + wr);
                    }
                }
            }
            QL += gsp * yl;
            QR += gsp * yr;
            QdL += gsp * (yr - yl);
            // rho_L and rho_R enter the face density alike
            if (sdiff && !ghost) rhob += gsp * (-ph.diff * 0.25 * area) * dot;
        }
        W[64 * (NFIX + q)] = make_double2(gp0, gp1);
    }
    double ubL[NE], ubR[NE], cL = 0.0, cR = 0.0;
#pragma unroll
    for (int p = 0; p < NE; p++) {
        ubL[p] = 0.0;
        ubR[p] = 0.0;
    }
    if (ph.flux == ABL_FV_FLUX_AUSM || ph.flux == ABL_FV_FLUX_RUSANOV)
        face_vjp<DIM, RUS>(fL, fR, QL, QR, QdL, nh, area, ph, g, ubL, ubR, cL, cR);
    // viscous flux
    double tqbf[NT], velbf[DIM], wg[2][DIM];
#pragma unroll
    for (int q = 0; q < NT; q++) tqbf[q] = 0.0;
#pragma unroll
    for (int d = 0; d < DIM; d++) velbf[d] = 0.0;
    if (visc) {
        double vL[DIM], vR[DIM], vf[DIM], tqf[NT], normal[DIM];
        const double irL = 1.0 / uL[kRho], irR = 1.0 / uR[kRho];
        double keL = 0.0, keR = 0.0;
#pragma unroll
        for (int d = 0; d < DIM; d++) {
            vL[d] = uL[kRhoU + d] * irL;
            vR[d] = uR[kRhoU + d] * irR;
            keL += vL[d] * vL[d];
            keR += vR[d] * vR[d];
            vf[d] = 0.5 * vL[d] + 0.5 * vR[d];
            normal[d] = area * nh[d];
        }
        if (ghost) {
            double aL[DIM + 1], aG[DIM + 1];
            load_face_pair<DIM>(m.fwa, f, wg);
            aL[0] = (uL[kRhoE] * irL - 0.5 * keL) * ph.icv;
            aG[0] = (uR[kRhoE] * irR - 0.5 * keR) * ph.icv;
#pragma unroll
            for (int d = 0; d < DIM; d++) {
                aL[1 + d] = vL[d];
                aG[1 + d] = vR[d];
            }
            double gA[DIM + 1][DIM];
#pragma unroll
            for (int q = 0; q <= DIM; q++)
#pragma unroll
                for (int d = 0; d < DIM; d++) gA[q][d] = wg[1][d] * (aL[q] - aG[q]);
            stress_heat<double, DIM>(ph.mu, ph.k, gA, tqR);
        }
#pragma unroll
        for (int q = 0; q < NT; q++) tqf[q] = 0.5 * tqL[q] + 0.5 * tqR[q];
        viscous_flux_vjp<DIM>(normal, vf, tqf, g, tqbf, velbf);
    }
    double tqh[NT], velh[DIM];
#pragma unroll
    for (int q = 0; q < NT; q++) tqh[q] = 0.5 * tqbf[q];
#pragma unroll
    for (int d = 0; d < DIM; d++) velh[d] = 0.5 * velbf[d];
    // ghost side (a ghost is always the right cell of its boundary face); the left cell's extras replace ubR
    double ugb[NE], extra[NE];
#pragma unroll
    for (int p = 0; p < NE; p++) {
        ugb[p] = ubR[p];
        extra[p] = 0.0;
    }
    if (ghost && visc) {
        double gAb[DIM + 1][DIM], aGb[DIM + 1];
#pragma unroll
        for (int q = 0; q <= DIM; q++) {
            aGb[q] = 0.0;
#pragma unroll
            for (int d = 0; d < DIM; d++) gAb[q][d] = 0.0;
        }
        stress_heat_vjp<DIM>(ph.mu, ph.k, tqh, gAb);
#pragma unroll
        for (int q = 0; q <= DIM; q++) {
            double t = 0.0;
#pragma unroll
            for (int d = 0; d < DIM; d++) t += wg[1][d] * gAb[q][d];
            extra[q] = t;  // a_L: the left cell's aux adjoint
            aGb[q] -= t;
        }
#pragma unroll
        for (int d = 0; d < DIM; d++) aGb[1 + d] += velh[d];
        aux_vjp<DIM>(uR, ph, aGb, ugb);
    }
    // species on boundary faces: the ghost's advection and diffusion adjoints (ghost gradients, face density)
    const double cdf = sdiff ? -ph.diff * (0.5 * uL[kRho] + 0.5 * uR[kRho]) * 0.5 * area : 0.0;
    double rgb = 0.0;
    for (int sp = 0; ghost && sp < ns; sp++) {
        const int64_t comp = NE + sp;
        const double2 gv = W[64 * (NFIX + (sp >> 1))];
        const double gsp = (sp & 1) ? gv.y : gv.x;
        double ygb = cR * gsp, yxl = 0.0;
        if (sdiff) {
            const double rl = uL[kRho], rr = uR[kRho];
            const double yl = U[comp * N + L], yr = U[comp * N + Rc];
            const double YL = yl / rl, YR = yr / rr;
            double dot = 0.0;
#pragma unroll
            for (int d = 0; d < DIM; d++) {
                const double GL = gS[sgp_word<DIM>(N, sp, L, DIM + d)];
                const double GR = wg[1][d] * (YL - YR);
                dot += nh[d] * (GL + GR);
            }
            const double drho = gsp * (-ph.diff * 0.25 * area) * dot;
            rhob += drho;  // rho_L and rho_G enter the face density alike
            // GY_ghost = w (Y_L - Y_G): the ghost's gradient adjoint goes to Y_L (left cell) and Y_G
            double t = 0.0;
#pragma unroll
            for (int d = 0; d < DIM; d++) t += wg[1][d] * gsp * cdf * nh[d];
            massfrac_vjp(rl, yl, t, yxl, extra[DIM + 1]);
            double rg = 0.0;
            massfrac_vjp(rr, yr, -t, ygb, rg);
            rgb += rg + drho;  // rho_G in the face density
        }
        b.Ughost[comp * NG + (Rc - NI)] = ygb;
        b.GL[(int64_t)sp * NG + (Rc - NI)] = yxl;
    }
    if (ghost) {
        ugb[kRho] += rgb;
#pragma unroll
        for (int p = 0; p < NE; p++) b.Ughost[p * NG + (Rc - NI)] = ugb[p];
    }
    // the record, pair by pair
    double w[AW::GS];
    w[AW::GS - 1] = 0.0;  // padding word (2-D)
#pragma unroll
    for (int p = 0; p < NE; p++) {
        w[AW::UL + p] = ubL[p];
        w[AW::UR + p] = ghost ? extra[p] : ubR[p];
    }
#pragma unroll
    for (int q = 0; q < NT; q++) w[AW::TQ + q] = tqh[q];
#pragma unroll
    for (int d = 0; d < DIM; d++) w[AW::VL + d] = velh[d];
    w[AW::CL] = cL;
    w[AW::CR] = cR;
    w[AW::CD] = cdf;
    w[AW::RH] = rhob;
#pragma unroll
    for (int q = 0; q < NFIX; q++) W[64 * q] = make_double2(w[2 * q], w[2 * q + 1]);
}

// Barth-Jespersen limiter of one component with its active pieces: phi = min(1, (umax - uc) / dpos, (umin - uc) /
// dneg); imax / imin = slot of the extreme neighbour (-1: the cell itself), kpos / kneg = slot of the extreme face
// projection.
struct BJFwd {
    double umax, umin, dpos, dneg, phi;
    int imax, imin, kpos, kneg, active;  // active: 0 none (phi = 1), 1 max side, 2 min side
};
__device__ inline void bj_finish(double uc, BJFwd& L) {
    L.phi = 1.0;
    L.active = 0;
    if (L.dpos > 0.0) {
        const double r = (L.umax - uc) / L.dpos;
        if (r < L.phi) {
            L.phi = r;
            L.active = 1;
        }
    }
    if (L.dneg < 0.0) {
        const double r = (L.umin - uc) / L.dneg;
        if (r < L.phi) {
            L.phi = r;
            L.active = 2;
        }
    }
}

// Reverse of one component's limited least-squares gradient g = phi(G) G, G = sum_k w_k (u_n - u_c), given the
// adjoint gbl of g: adds the adjoint of the cell's own value to ubc and of each neighbour's to cn[k] (ghost
// neighbours through to_ghost).  unv[k] = the component's value in the neighbour across slot k, x[k] = x_f - x_c of
// slot k (the Barth-Jespersen projection vector).
template <int DIM, int KM, typename ToGhost>
__device__ inline void limited_lsq_vjp(const double* unv, double uc, const double (*w)[DIM], const double (*x)[DIM],
                                       const int64_t* nb, const bool* valid, const bool* bnd, bool bj,
                                       const double* gbl, double& ubc, double* cn, ToGhost&& to_ghost) {
    double G[DIM], Gb[DIM];
#pragma unroll
    for (int d = 0; d < DIM; d++) G[d] = 0.0;
    BJFwd lim{uc, uc, 0.0, 0.0, 1.0, -1, -1, -1, -1, 0};
#pragma unroll
    for (int k = 0; k < KM; k++) {
        if (!valid[k]) continue;
        const double un = unv[k];
        if (!bnd[k])
#pragma unroll
            for (int d = 0; d < DIM; d++) G[d] += w[k][d] * (un - uc);
        if (un > lim.umax) {
            lim.umax = un;
            lim.imax = k;
        }
        if (un < lim.umin) {
            lim.umin = un;
            lim.imin = k;
        }
    }
    if (bj) {
#pragma unroll
        for (int k = 0; k < KM; k++) {
            if (!valid[k]) continue;
            double D = 0.0;
#pragma unroll
            for (int d = 0; d < DIM; d++) D += G[d] * x[k][d];
            if (D > lim.dpos) {
                lim.dpos = D;
                lim.kpos = k;
            }
            if (D < lim.dneg) {
                lim.dneg = D;
                lim.kneg = k;
            }
        }
        bj_finish(uc, lim);
    }
    double phib = 0.0;
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        Gb[d] = lim.phi * gbl[d];
        phib += gbl[d] * G[d];
    }
    if (bj && lim.active) {
        const bool mx = lim.active == 1;
        const double den = mx ? lim.dpos : lim.dneg;
        const int ext = mx ? lim.imax : lim.imin;
        const int kd = mx ? lim.kpos : lim.kneg;
        const double ub_ext = phib / den;
        ubc -= ub_ext;
        if (ext < 0) ubc += ub_ext;
        const double den_b = -lim.phi * phib / den;
#pragma unroll
        for (int k = 0; k < KM; k++) {  // selects, not runtime register indices (which would spill)
            if (k == ext) {
                if (bnd[k])
                    to_ghost(nb[k], ub_ext);
                else
                    cn[k] += ub_ext;
            }
            if (k == kd)
#pragma unroll
                for (int d = 0; d < DIM; d++) Gb[d] += den_b * x[k][d];
        }
    }
    // least-squares gradient: G = sum_k w_k (u_n - u_c)
#pragma unroll
    for (int k = 0; k < KM; k++) {
        if (!valid[k] || bnd[k]) continue;
        double t = 0.0;
#pragma unroll
        for (int d = 0; d < DIM; d++) t += w[k][d] * Gb[d];
        cn[k] += t;
        ubc -= t;
    }
}

// Pass 2, euler field (lane per interior cell): sums its faces' adjoint records in slot order (state, limited
// gradients through the face state U + G dx, viscous terms, face velocity and density, boundary-face extras), then
// reverses the limiter, the least-squares gradients and the aux (temperature / velocity) gradients of the cell;
// writes its own adjoint and, per face slot, the contribution for the neighbour across the face.
template <int DIM, int KM>
__global__ __launch_bounds__(kBlock) void k_adj_cells_e(MeshView<double> m, Phys ph, const double* __restrict__ U,
                                                        AdjBufs b) {
    constexpr int NE = DIM + 2;
    constexpr int NT = ViscTerms<DIM>::NT;
    using AW = AdjFaceRec<DIM>;
    const int64_t c = xcd_cell_index();
    if (c >= m.NI) return;
    const int64_t N = m.N, NI = m.NI, NG = N - NI;
    const int nc = NE + ph.ns;
    const bool recon = ph.recon != 0, visc = ph.visc != 0;
    const bool bj = ph.limiter == ABL_FV_LIMITER_BARTH_JESPERSEN;
    const double2* __restrict__ W = reinterpret_cast<const double2*>(b.W);
    double ub[NE], gb[NE][DIM], tqb[NT], velb[DIM], auxb[DIM + 1], rhob = 0.0;
#pragma unroll
    for (int p = 0; p < NE; p++) {
        ub[p] = 0.0;
#pragma unroll
        for (int d = 0; d < DIM; d++) gb[p][d] = 0.0;
    }
#pragma unroll
    for (int q = 0; q < NT; q++) tqb[q] = 0.0;
#pragma unroll
    for (int d = 0; d < DIM; d++) velb[d] = 0.0;
#pragma unroll
    for (int q = 0; q <= DIM; q++) auxb[q] = 0.0;
    int64_t nb[KM];
    bool valid[KM], bnd[KM];
    double xs[KM][DIM];
#pragma unroll
    for (int k = 0; k < KM; k++) {
        const int32_t e = k < m.K ? m.ell[k * NI + c] : -1;
        valid[k] = e >= 0;
        const int64_t f = valid[k] ? (int64_t)(e >> 1) : 0;
        const int s = e & 1;
        const int32_t n = m.snbr[k * NI + c];
        nb[k] = valid[k] ? (int64_t)n : c;
        bnd[k] = nb[k] >= NI;
        double x[2][DIM];
        load_face_pair<DIM>(m.fdx, f, x);
#pragma unroll
        for (int d = 0; d < DIM; d++) xs[k][d] = s ? x[1][d] : x[0][d];
        // the record pairs this cell needs: its side's face-state adjoint, the viscous terms and face velocity, the
        // face density adjoint, and on a boundary face the right-side words (the left cell's extras)
        constexpr int NPS = (NE + 1) / 2;                      // pairs covering NE words from either parity
        constexpr int V0 = AW::TQ / 2, V1 = (AW::VL + DIM - 1) / 2;
        const int64_t wb = adjw_pair(f, b.nwp);
        const int o = s ? AW::UR : AW::UL;
        const int p0 = o / 2;
        double sw_[2 * NPS], vw[2 * (V1 - V0 + 1)];
#pragma unroll
        for (int i = 0; i < NPS; i++) {
            const double2 v = W[wb + 64 * (p0 + i)];
            sw_[2 * i] = v.x;
            sw_[2 * i + 1] = v.y;
        }
        if (visc)
#pragma unroll
            for (int q = V0; q <= V1; q++) {
                const double2 v = W[wb + 64 * q];
                vw[2 * (q - V0)] = v.x;
                vw[2 * (q - V0) + 1] = v.y;
            }
        const double2 rp = W[wb + 64 * (AW::RH / 2)];
        const double rh = (AW::RH & 1) ? rp.y : rp.x;
        if (!valid[k]) continue;
#pragma unroll
        for (int p = 0; p < NE; p++) {
            const double wp = sw_[o + p - 2 * p0];
            ub[p] += wp;
            if (recon)
#pragma unroll
                for (int d = 0; d < DIM; d++) gb[p][d] += wp * xs[k][d];
        }
        if (visc) {
#pragma unroll
            for (int q = 0; q < NT; q++) tqb[q] += vw[AW::TQ + q - 2 * V0];
#pragma unroll
            for (int d = 0; d < DIM; d++) velb[d] += vw[AW::VL + d - 2 * V0];
        }
        rhob += rh;
        if (bnd[k]) {  // boundary face: the left cell's extras from the ghost's viscous gradient
            constexpr int R0 = AW::UR / 2;
            double rw[2 * NPS];
#pragma unroll
            for (int i = 0; i < NPS; i++) {
                const double2 v = W[wb + 64 * (R0 + i)];
                rw[2 * i] = v.x;
                rw[2 * i + 1] = v.y;
            }
#pragma unroll
            for (int q = 0; q <= DIM; q++) auxb[q] += rw[AW::UR + q - 2 * R0];
            rhob += rw[AW::UR + DIM + 1 - 2 * R0];
        }
    }
    double uc[NE];
#pragma unroll
    for (int p = 0; p < NE; p++) uc[p] = U[p * N + c];
    if (visc) {
        double ab[DIM + 1];
        ab[0] = auxb[0];
#pragma unroll
        for (int d = 0; d < DIM; d++) ab[1 + d] = velb[d] + auxb[1 + d];
        aux_vjp<DIM>(uc, ph, ab, ub);
    }
    ub[kRho] += rhob;
    // the cell's reverse: limiter and least-squares gradients, then the aux gradients
    double ce[KM][NE];
#pragma unroll
    for (int k = 0; k < KM; k++)
#pragma unroll
        for (int p = 0; p < NE; p++) ce[k][p] = 0.0;
    double unb_all[KM][NE];  // neighbour states (the limiter / least-squares and the aux reverses both read them)
#pragma unroll
    for (int k = 0; k < KM; k++)
#pragma unroll
        for (int p = 0; p < NE; p++) unb_all[k][p] = U[p * N + nb[k]];
    if (recon) {
        double wl[KM][DIM];
#pragma unroll
        for (int k = 0; k < KM; k++)
#pragma unroll
            for (int d = 0; d < DIM; d++) wl[k][d] = m.sw[(k * DIM + d) * NI + c];
#pragma unroll
        for (int p = 0; p < NE; p++) {
            double cn[KM], unv[KM];
#pragma unroll
            for (int k = 0; k < KM; k++) {
                cn[k] = 0.0;
                unv[k] = unb_all[k][p];
            }
            limited_lsq_vjp<DIM, KM>(unv, uc[p], wl, xs, nb, valid, bnd, bj, gb[p], ub[p], cn,
                                     [&](int64_t g, double v) { b.Ughost[(int64_t)p * NG + (g - NI)] += v; });
#pragma unroll
            for (int k = 0; k < KM; k++) ce[k][p] += cn[k];
        }
    }
    if (visc) {
        const double* __restrict__ wa_arr = m.swa ? m.swa : m.sw;
        double gAb[DIM + 1][DIM], acb[DIM + 1];
#pragma unroll
        for (int q = 0; q <= DIM; q++) {
            acb[q] = 0.0;
#pragma unroll
            for (int d = 0; d < DIM; d++) gAb[q][d] = 0.0;
        }
        stress_heat_vjp<DIM>(ph.mu, ph.k, tqb, gAb);
#pragma unroll
        for (int k = 0; k < KM; k++) {
            if (!valid[k]) continue;
            double anb[DIM + 1], un[NE], unb[NE];
#pragma unroll
            for (int q = 0; q <= DIM; q++) {
                double t = 0.0;
#pragma unroll
                for (int d = 0; d < DIM; d++) t += wa_arr[(k * DIM + d) * NI + c] * gAb[q][d];
                anb[q] = t;
                acb[q] -= t;
            }
#pragma unroll
            for (int p = 0; p < NE; p++) {
                un[p] = unb_all[k][p];
                unb[p] = 0.0;
            }
            aux_vjp<DIM>(un, ph, anb, unb);
#pragma unroll
            for (int p = 0; p < NE; p++) {
                if (bnd[k])
                    b.Ughost[(int64_t)p * NG + (nb[k] - NI)] += unb[p];
                else
                    ce[k][p] += unb[p];
            }
        }
        aux_vjp<DIM>(uc, ph, acb, ub);
    }
#pragma unroll
    for (int p = 0; p < NE; p++) b.Uself[(int64_t)p * NI + c] = ub[p];
#pragma unroll
    for (int k = 0; k < KM; k++)
#pragma unroll
        for (int p = 0; p < NE; p++) b.contrib[((int64_t)k * nc + p) * NI + c] = ce[k][p];
}

// Pass 2, species (lane per interior cell, after k_adj_cells_e): per species, sums the species factors of its faces'
// records (densityYi adjoint, limited densityYi gradient adjoint, Yi aux gradient adjoint), then reverses the
// limiter, the least-squares gradient and the mass-fraction gradient.  The slot geometry stays in registers across
// the species loop.  The species' density terms (Y = rhoY / rho) are added into the euler kernel's density rows.
template <int DIM, int KM>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(2))) void k_adj_cells_s(MeshView<double> m, Phys ph, const double* __restrict__ U,
                                                        AdjBufs b) {
    constexpr int NE = DIM + 2;
    using AW = AdjFaceRec<DIM>;
    const int64_t c = xcd_cell_index();
    if (c >= m.NI) return;
    const int64_t N = m.N, NI = m.NI, NG = N - NI;
    const int ns = ph.ns, nc = NE + ns;
    const bool recon = ph.recon != 0, sdiff = ph.sdiff != 0;
    const bool bj = ph.limiter == ABL_FV_LIMITER_BARTH_JESPERSEN;
    const double2* __restrict__ W = reinterpret_cast<const double2*>(b.W);
    const double* __restrict__ wa_arr = m.swa ? m.swa : m.sw;
    int64_t nb[KM], wb[KM], gi[KM];
    bool valid[KM], bnd[KM];
    double xs[KM][DIM], cnv[KM][DIM], csk[KM], wl[KM][DIM], wa[KM][DIM];
#pragma unroll
    for (int k = 0; k < KM; k++) {
        const int32_t e = k < m.K ? m.ell[k * NI + c] : -1;
        valid[k] = e >= 0;
        const int64_t f = valid[k] ? (int64_t)(e >> 1) : 0;
        const int s = e & 1;
        const int32_t n = m.snbr[k * NI + c];
        nb[k] = valid[k] ? (int64_t)n : c;
        bnd[k] = nb[k] >= NI;
        gi[k] = (valid[k] && bnd[k]) ? nb[k] - NI : -1;
        wb[k] = adjw_pair(f, b.nwp);
        double x[2][DIM], nh[3], area;
        load_face_pair<DIM>(m.fdx, f, x);
        load_face_na(m.fna, f, nh, area);
        double cw[4];  // words CL, CR, CD of the record (two pairs from CL's pair on)
        {
            const double2 v0 = W[wb[k] + 64 * (AW::CL / 2)], v1 = W[wb[k] + 64 * (AW::CL / 2 + 1)];
            cw[0] = v0.x;
            cw[1] = v0.y;
            cw[2] = v1.x;
            cw[3] = v1.y;
        }
        const int o = AW::CL & 1;
        csk[k] = valid[k] ? (s ? cw[o + 1] : cw[o]) : 0.0;
        const double cd = valid[k] ? cw[o + 2] : 0.0;
#pragma unroll
        for (int d = 0; d < DIM; d++) {
            xs[k][d] = s ? x[1][d] : x[0][d];
            cnv[k][d] = cd * nh[d];
            wl[k][d] = recon ? m.sw[(k * DIM + d) * NI + c] : 0.0;
            wa[k][d] = sdiff ? wa_arr[(k * DIM + d) * NI + c] : 0.0;
        }
    }
    const double rho_c = U[c];
    double rbc = 0.0, rbn[KM], rho_n[KM];
#pragma unroll
    for (int k = 0; k < KM; k++) {
        rbn[k] = 0.0;
        rho_n[k] = sdiff ? U[nb[k]] : 1.0;
    }
    const double* __restrict__ GL = b.GL;
    // species in pairs: both species' loads (own and neighbour densities, the record pair) are issued before either
    // is reversed, so two dependent rounds of memory latency overlap
    for (int sp0 = 0; sp0 < ns; sp0 += 2) {
        const int nh = sp0 + 1 < ns ? 2 : 1;
        double ycs[2], unv[2][KM], gsp[2][KM];
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int64_t comp = NE + sp0 + (h < nh ? h : 0);
            ycs[h] = U[comp * N + c];
#pragma unroll
            for (int k = 0; k < KM; k++) unv[h][k] = U[comp * N + nb[k]];
        }
#pragma unroll
        for (int k = 0; k < KM; k++) {
            const double2 v = W[wb[k] + 64 * (AW::GS / 2 + (sp0 >> 1))];
            gsp[0][k] = v.x;
            gsp[1][k] = v.y;
        }
#pragma unroll
        for (int h = 0; h < 2; h++) {
            if (h >= nh) break;
            const int sp = sp0 + h;
            const int64_t comp = NE + sp;
            // the faces' species factors
            double usb = 0.0, gsb[DIM], gyb[DIM];
#pragma unroll
            for (int d = 0; d < DIM; d++) {
                gsb[d] = 0.0;
                gyb[d] = 0.0;
            }
#pragma unroll
            for (int k = 0; k < KM; k++) {
                if (!valid[k]) continue;
                const double yb = csk[k] * gsp[h][k];
                usb += yb;
                if (gi[k] >= 0) usb += GL[(int64_t)sp * NG + gi[k]];
#pragma unroll
                for (int d = 0; d < DIM; d++) {
                    gsb[d] += yb * xs[k][d];
                    gyb[d] += gsp[h][k] * cnv[k][d];
                }
            }
            double cs[KM];  // this species' contributions to the neighbours
#pragma unroll
            for (int k = 0; k < KM; k++) cs[k] = 0.0;
            if (recon)
                limited_lsq_vjp<DIM, KM>(unv[h], ycs[h], wl, xs, nb, valid, bnd, bj, gsb, usb, cs,
                                         [&](int64_t g, double v) { b.Ughost[comp * NG + (g - NI)] += v; });
            if (sdiff) {
                double ycb = 0.0;
#pragma unroll
                for (int k = 0; k < KM; k++) {
                    if (!valid[k]) continue;
                    double t = 0.0;
#pragma unroll
                    for (int d = 0; d < DIM; d++) t += wa[k][d] * gyb[d];
                    ycb -= t;
                    double yb_n = 0.0, rb_n = 0.0;
                    massfrac_vjp(rho_n[k], unv[h][k], t, yb_n, rb_n);
                    if (bnd[k]) {
                        b.Ughost[comp * NG + (nb[k] - NI)] += yb_n;
                        b.Ughost[(int64_t)kRho * NG + (nb[k] - NI)] += rb_n;
                    } else {
                        cs[k] += yb_n;
                        rbn[k] += rb_n;
                    }
                }
                massfrac_vjp(rho_c, ycs[h], ycb, usb, rbc);
            }
            b.Uself[comp * NI + c] = usb;
#pragma unroll
            for (int k = 0; k < KM; k++) b.contrib[((int64_t)k * nc + comp) * NI + c] = cs[k];
        }
    }
    if (sdiff) {
        b.Uself[(int64_t)kRho * NI + c] += rbc;
#pragma unroll
        for (int k = 0; k < KM; k++)
            if (valid[k] && !bnd[k]) b.contrib[((int64_t)k * nc + kRho) * NI + c] += rbn[k];
    }
}

template <int DIM, int KM>
__global__ __launch_bounds__(kBlock) void k_adj_gather(MeshView<double> m, Phys ph,
                                                       const int32_t* __restrict__ ghost_slot,
                                                       const int32_t* __restrict__ slot_kind, AdjBufs b,
                                                       double* __restrict__ Ub) {
    constexpr int NE = DIM + 2;
    const int64_t c = xcd_cell_index();
    if (c >= m.NI) return;
    const int64_t N = m.N, NI = m.NI, NG = N - NI;
    const int nc = NE + ph.ns;
    int64_t nb[KM];
    int mk[KM];
    double from_ghost[KM];  // dU_ghost / dU_cell: +1 (OUTFLOW copy), -1 (enforceAtFace inflow), 0 otherwise
#pragma unroll
    for (int k = 0; k < KM; k++) {
        const int32_t e = m.snbr[k * NI + c];
        nb[k] = e;
        mk[k] = b.mirror[k * NI + c];
        const int gk = e >= NI ? slot_kind[ghost_slot[e - NI]] : ABL_FV_BC_NONE;
        from_ghost[k] = gk == ABL_FV_BC_OUTFLOW ? 1.0 : (gk == kBcInflowAtFace ? -1.0 : 0.0);
    }
    for (int p = 0; p < nc; p++) {
        double acc = b.Uself[(int64_t)p * NI + c];
#pragma unroll
        for (int k = 0; k < KM; k++) {
            if (nb[k] < 0) continue;
            if (nb[k] < NI)
                acc += b.contrib[((int64_t)mk[k] * nc + p) * NI + nb[k]];
            else if (from_ghost[k] != 0.0)
                acc += from_ghost[k] * b.Ughost[(int64_t)p * NG + (nb[k] - NI)];
        }
        Ub[(int64_t)p * N + c] = acc;
    }
}

// Adjoint rows of caller-filled ghost cells (ABL_FV_BC_GHOST): their values are inputs of the RHS, so their
// accumulated adjoint is returned in the ghost rows of Ubar (the other ghost rows were zeroed).
__global__ __launch_bounds__(kBlock) void k_ghost_adjoint_out(int64_t NG, int64_t N, int64_t NI, int nc,
                                                              const int32_t* __restrict__ ghost_slot,
                                                              const int32_t* __restrict__ slot_kind,
                                                              const double* __restrict__ Ughost, double* __restrict__ Ub) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= NG || slot_kind[ghost_slot[g]] != ABL_FV_BC_GHOST) return;
    for (int p = 0; p < nc; p++) Ub[(int64_t)p * N + NI + g] = Ughost[(int64_t)p * NG + g];
}

// Per-slot sums of the per-boundary-face theta_bar contributions: block 3 s + j sums component j over the faces of
// slot s, each lane over a fixed stride of faces with four loads in flight, then a fixed-order tree
// (deterministic).  Only inflow slots have nonzero contributions; the blocks of the other slots write zero.
constexpr int kThetaBlock = 1024;
__global__ __launch_bounds__(kThetaBlock) void k_theta_reduce(int64_t nb, const int32_t* __restrict__ bslot,
                                                              const int32_t* __restrict__ slot_kind,
                                                              const double* __restrict__ contrib,
                                                              double* __restrict__ out) {
    const int s = blockIdx.x / 3, j = blockIdx.x % 3;
    const int kind = slot_kind[s];
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    if (kind == ABL_FV_BC_INFLOW || kind == kBcInflowAtFace) {
        int64_t i = threadIdx.x;
        for (; i + 3 * kThetaBlock < nb; i += 4 * kThetaBlock) {
            const int32_t b0 = bslot[i], b1 = bslot[i + kThetaBlock], b2 = bslot[i + 2 * kThetaBlock],
                          b3 = bslot[i + 3 * kThetaBlock];
            if (b0 == s) a0 += contrib[3 * i + j];
            if (b1 == s) a1 += contrib[3 * (i + kThetaBlock) + j];
            if (b2 == s) a2 += contrib[3 * (i + 2 * kThetaBlock) + j];
            if (b3 == s) a3 += contrib[3 * (i + 3 * kThetaBlock) + j];
        }
        for (; i < nb; i += kThetaBlock)
            if (bslot[i] == s) a0 += contrib[3 * i + j];
    }
    __shared__ double red[kThetaBlock];
    red[threadIdx.x] = (a0 + a1) + (a2 + a3);
    __syncthreads();
    for (int w = kThetaBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

// theta_bar contributions of each inflow boundary face from its ghost's adjoint: dU_ghost/dtheta^T Ubar_ghost.
__global__ __launch_bounds__(kBlock) void k_adj_theta(int64_t nb, int dim, int ns, int64_t N, int64_t NI,
                                                      const int32_t* __restrict__ bface, const int32_t* __restrict__ bslot,
                                                      const int32_t* __restrict__ fcell, int64_t F,
                                                      const int32_t* __restrict__ slot_kind,
                                                      const double* __restrict__ slot_param,
                                                      const double* __restrict__ slot_Y, Phys ph,
                                                      const double* __restrict__ Ughost, double* __restrict__ contrib) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nb) return;
    const int slot = bslot[i];
    double t0 = 0.0, t1 = 0.0, t2 = 0.0;
    const int kind = slot_kind[slot];
    if (kind == ABL_FV_BC_INFLOW || kind == kBcInflowAtFace) {
        const int64_t g = fcell[F + bface[i]] - NI, NG = N - NI;
        const double* sp_ = slot_param + 6 * slot;
        const double rho = sp_[0], um = sp_[1], T = sp_[2];
        const double ubr = Ughost[kRho * NG + g], ube = Ughost[kRhoE * NG + g];
        t0 = ubr + ube * (ph.cv * T + 0.5 * um * um);
        t1 = ube * rho * um;
        t2 = ube * rho * ph.cv;
        for (int d = 0; d < dim; d++) {
            const double ubm = Ughost[(kRhoU + d) * NG + g];
            t0 += ubm * um * sp_[3 + d];
            t1 += ubm * rho * sp_[3 + d];
        }
        for (int sp = 0; sp < ns; sp++) t0 += Ughost[(dim + 2 + sp) * NG + g] * slot_Y[slot * ns + sp];
        if (kind == kBcInflowAtFace) {  // ghost = 2 state(theta) - interior
            t0 *= 2.0;
            t1 *= 2.0;
            t2 *= 2.0;
        }
    }
    contrib[3 * i + 0] = t0;
    contrib[3 * i + 1] = t1;
    contrib[3 * i + 2] = t2;
}

// FaceInterpolant::GetInterpolatedFaceVectors (faceInterpolant.cpp:109-249) with the LeastSquaresAverage stencils
// (stencils/leastSquaresAverage.cpp:5-155), one lane per face: value 1/2 (q_L + q_R) and gradient 1/2 (G_L + G_R) of
// every solution component and every aux field (T, velocity, Yi), G_c the unlimited least-squares gradient over all
// face neighbours of c (the face-stencil weight set the viscous fluxes use) and, for a ghost, its one-neighbour
// gradient w_g (q_interior - q_ghost).  Outputs [q][F] and [q * DIM + d][F]; null outputs are skipped.
template <int DIM, int KM>
__device__ inline double face_interp_value(const double* __restrict__ U, int64_t N, int nc, const Phys& ph, int q,
                                           int64_t c) {
    if (q < nc) return U[(int64_t)q * N + c];
    const double rho = U[c], irho = 1.0 / rho;
    const int a = q - nc;
    if (a == 0) {  // PerfectGas temperature (perfectGas.cpp:76-86)
        double ke = 0.0;
#pragma unroll
        for (int d = 0; d < DIM; d++) {
            const double v = U[(int64_t)(kRhoU + d) * N + c] * irho;
            ke += v * v;
        }
        return (U[(int64_t)kRhoE * N + c] * irho - 0.5 * ke) * ph.icv;
    }
    if (a <= DIM) return U[(int64_t)(kRhoU + a - 1) * N + c] * irho;
    return U[(int64_t)(DIM + 2 + a - 1 - DIM) * N + c] * irho;
}
template <int DIM, int KM>
__global__ __launch_bounds__(kBlock) void k_face_interpolate(MeshView<double> m, Phys ph, const double* __restrict__ U,
                                                             int nc, double* __restrict__ fsol,
                                                             double* __restrict__ fsolg, double* __restrict__ faux,
                                                             double* __restrict__ fauxg) {
    const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= m.F) return;
    const int64_t N = m.N, NI = m.NI, F = m.F;
    const int64_t L = m.fcell[f], R = m.fcell[F + f];
    const bool ghost = R >= NI;
    const double* __restrict__ wa_arr = m.swa ? m.swa : m.sw;
    const int nq = nc + 1 + DIM + ph.ns;
    for (int q = 0; q < nq; q++) {
        const double vL = face_interp_value<DIM, KM>(U, N, nc, ph, q, L);
        const double vR = face_interp_value<DIM, KM>(U, N, nc, ph, q, R);
        double g[DIM];
#pragma unroll
        for (int d = 0; d < DIM; d++) g[d] = 0.0;
        for (int sd = 0; sd < 2; sd++) {
            const int64_t c = sd == 0 ? L : R;
            const double vc = sd == 0 ? vL : vR;
            if (sd == 1 && ghost) {
#pragma unroll
                for (int d = 0; d < DIM; d++) g[d] += 0.5 * (m.fwa[2 * DIM * f + DIM + d] * (vL - vR));
                continue;
            }
            double gc[DIM];
#pragma unroll
            for (int d = 0; d < DIM; d++) gc[d] = 0.0;
            for (int k = 0; k < KM; k++) {
                const int32_t e = m.snbr[k * NI + c];
                if (e < 0) continue;
                const double dv = face_interp_value<DIM, KM>(U, N, nc, ph, q, e) - vc;
#pragma unroll
                for (int d = 0; d < DIM; d++) gc[d] += wa_arr[(k * DIM + d) * NI + c] * dv;
            }
#pragma unroll
            for (int d = 0; d < DIM; d++) g[d] += 0.5 * gc[d];
        }
        double* val = q < nc ? fsol : faux;
        double* grd = q < nc ? fsolg : fauxg;
        const int qq = q < nc ? q : q - nc;
        if (val) val[(int64_t)qq * F + f] = 0.5 * vL + 0.5 * vR;
        if (grd)
#pragma unroll
            for (int d = 0; d < DIM; d++) grd[((int64_t)qq * DIM + d) * F + f] = g[d];
    }
}
using FaceInterpKernel = void (*)(MeshView<double>, Phys, const double*, int, double*, double*, double*, double*);
template <int DIM>
FaceInterpKernel pick_face_interpolate(int K) {
    switch (K <= 4 ? 4 : (K <= 6 ? 6 : 8)) {
        case 4: return k_face_interpolate<DIM, 4>;
        case 6: return k_face_interpolate<DIM, 6>;
        default: return k_face_interpolate<DIM, 8>;
    }
}

// ---- compile-time variant selection
using AdjCellsKernel = void (*)(MeshView<double>, Phys, const double*, AdjBufs);
using AdjFacesKernel = void (*)(MeshView<double>, Phys, const double*, const double*, const double*, const double*,
                                const double*, AdjBufs);
using AdjGatherKernel = void (*)(MeshView<double>, Phys, const int32_t*, const int32_t*, AdjBufs, double*);
template <int DIM, int KM>
void pick_adj_k(bool rusanov, AdjFacesKernel& fk, AdjCellsKernel& ce, AdjCellsKernel& cs, AdjGatherKernel& gk) {
    fk = rusanov ? k_adj_face_rev<DIM, true> : k_adj_face_rev<DIM, false>;
    ce = k_adj_cells_e<DIM, KM>;
    cs = k_adj_cells_s<DIM, KM>;
    gk = k_adj_gather<DIM, KM>;
}
template <typename G>
using GradKernelT = void (*)(MeshView<G>, Phys, const double*, G*, G*, void*);
using GradKernel = GradKernelT<double>;
template <typename G>
using SpeciesGradKernelT = void (*)(MeshView<G>, Phys, const double*, G*, G*);
template <typename G, int DIM>
SpeciesGradKernelT<G> pick_species_gradients(int K) {
    switch (K <= 4 ? 4 : (K <= 6 ? 6 : 8)) {
        case 4: return k_species_gradients_w3<G, DIM, 4>;
        case 6: return k_species_gradients<G, DIM, 6>;
        default: return k_species_gradients<G, DIM, 8>;
    }
}
using FluxKernel = void (*)(MeshView<double>, Phys, const double*, const double*, const double*, const double*,
                            double*, int);
using FaceFluxKernel = void (*)(MeshView<double>, Phys, const double*, const double*, const double*, const double*,
                                double*, double*);
using GatherKernel = void (*)(MeshView<double>, int, const double*, const double*, double*, int);

template <typename G, int DIM, int KM>
GradKernelT<G> pick_gradients_k(bool recon, bool visc, int lim) {
    if (!recon) return visc ? k_gradients<G, DIM, false, true, kLimNone, KM> : nullptr;
    if constexpr (sizeof(G) == 8 && DIM == 3) {
        if (lim != kLimPetsc) {
            if (visc) return lim == kLimBJ ? k_gradients_w4<G, DIM, true, true, kLimBJ, KM>
                                           : k_gradients_w4<G, DIM, true, true, kLimNone, KM>;
            return lim == kLimBJ ? k_gradients_w4<G, DIM, true, false, kLimBJ, KM>
                                 : k_gradients_w4<G, DIM, true, false, kLimNone, KM>;
        }
    }
    if (visc) {
        if (lim == kLimBJ) return k_gradients<G, DIM, true, true, kLimBJ, KM>;
        if (lim == kLimPetsc) return k_gradients<G, DIM, true, true, kLimPetsc, KM>;
        return k_gradients<G, DIM, true, true, kLimNone, KM>;
    }
    if (lim == kLimBJ) return k_gradients<G, DIM, true, false, kLimBJ, KM>;
    if (lim == kLimPetsc) return k_gradients<G, DIM, true, false, kLimPetsc, KM>;
    return k_gradients<G, DIM, true, false, kLimNone, KM>;
}
// Compile-time slot width of the cell -> face table: 4 (quads, tets), 6 (hexes) or 8.
inline int slot_width(int K) { return K <= 4 ? 4 : (K <= 6 ? 6 : 8); }
template <int DIM>
void pick_adj(bool rusanov, int K, AdjFacesKernel& fk, AdjCellsKernel& ce, AdjCellsKernel& cs,
              AdjGatherKernel& gk) {
    switch (slot_width(K)) {
        case 4: pick_adj_k<DIM, 4>(rusanov, fk, ce, cs, gk); break;
        case 6: pick_adj_k<DIM, 6>(rusanov, fk, ce, cs, gk); break;
        default: pick_adj_k<DIM, 8>(rusanov, fk, ce, cs, gk); break;
    }
}
template <typename G, int DIM>
GradKernelT<G> pick_gradients(bool recon, bool visc, int lim, int K) {
    switch (slot_width(K)) {
        case 4: return pick_gradients_k<G, DIM, 4>(recon, visc, lim);
        case 6: return pick_gradients_k<G, DIM, 6>(recon, visc, lim);
        default: return pick_gradients_k<G, DIM, 8>(recon, visc, lim);
    }
}

template <int DIM, bool RECON>
FluxKernel pick_flux2(bool visc, bool species) {
    if (visc) return species ? k_flux<DIM, RECON, true, true> : k_flux<DIM, RECON, true, false>;
    return species ? k_flux<DIM, RECON, false, true> : k_flux<DIM, RECON, false, false>;
}
template <int DIM>
FluxKernel pick_flux(bool recon, bool visc, bool species) {
    return recon ? pick_flux2<DIM, true>(visc, species) : pick_flux2<DIM, false>(visc, species);
}
template <int DIM, bool RECON>
FaceFluxKernel pick_face_flux2(bool visc, bool species) {
    if (visc) return species ? k_face_flux<DIM, RECON, true, true> : k_face_flux<DIM, RECON, true, false>;
    return species ? k_face_flux<DIM, RECON, false, true> : k_face_flux<DIM, RECON, false, false>;
}
template <int DIM>
FaceFluxKernel pick_face_flux(bool recon, bool visc, bool species) {
    return recon ? pick_face_flux2<DIM, true>(visc, species) : pick_face_flux2<DIM, false>(visc, species);
}
using MixedFluxKernel = void (*)(MeshView<float>, Phys, const double*, const float*, double*, int);
template <int DIM, bool RUS>
MixedFluxKernel pick_flux_mixed2(bool recon, bool visc) {
    if (recon) return visc ? k_flux_mixed<DIM, true, true, RUS> : k_flux_mixed<DIM, true, false, RUS>;
    return visc ? k_flux_mixed<DIM, false, true, RUS> : k_flux_mixed<DIM, false, false, RUS>;
}
template <int DIM>
MixedFluxKernel pick_flux_mixed(bool recon, bool visc, bool rus) {
    return rus ? pick_flux_mixed2<DIM, true>(recon, visc) : pick_flux_mixed2<DIM, false>(recon, visc);
}
using TileMixedKernel = void (*)(MeshView<float>, Phys, const double*, const float*, const float*, const float*,
                                 const int32_t*, const int32_t*, const int32_t*, int, double*, int);
template <int DIM, bool RECON, bool VISC, bool SPECIES, bool RUS>
TileMixedKernel pick_tile_mixed3(int K) {
    switch (slot_width(K)) {
        case 4: return k_tile_flux_mixed<DIM, RECON, VISC, SPECIES, 4, RUS>;
        case 6: return k_tile_flux_mixed<DIM, RECON, VISC, SPECIES, 6, RUS>;
        default: return k_tile_flux_mixed<DIM, RECON, VISC, SPECIES, 8, RUS>;
    }
}
template <int DIM, bool RECON, bool VISC, bool RUS>
TileMixedKernel pick_tile_mixed2(bool species, int K) {
    return species ? pick_tile_mixed3<DIM, RECON, VISC, true, RUS>(K) : pick_tile_mixed3<DIM, RECON, VISC, false, RUS>(K);
}
using FaceFluxMixedKernel = void (*)(MeshView<float>, Phys, const double*, const float*, const float*, const float*,
                                     float*, int);
template <int DIM, bool RECON, bool VISC, bool RUS>
FaceFluxMixedKernel pick_face_flux_mixed2(bool species) {
    return species ? k_face_flux_mixed<DIM, RECON, VISC, true, RUS> : k_face_flux_mixed<DIM, RECON, VISC, false, RUS>;
}
template <int DIM, bool RUS>
FaceFluxMixedKernel pick_face_flux_mixed1(bool recon, bool visc, bool species) {
    if (recon)
        return visc ? pick_face_flux_mixed2<DIM, true, true, RUS>(species)
                    : pick_face_flux_mixed2<DIM, true, false, RUS>(species);
    return visc ? pick_face_flux_mixed2<DIM, false, true, RUS>(species)
                : pick_face_flux_mixed2<DIM, false, false, RUS>(species);
}
template <int DIM>
FaceFluxMixedKernel pick_face_flux_mixed(bool recon, bool visc, bool species, bool rus) {
    return rus ? pick_face_flux_mixed1<DIM, true>(recon, visc, species)
               : pick_face_flux_mixed1<DIM, false>(recon, visc, species);
}
using GatherMixedKernel = void (*)(MeshView<float>, int, const float*, double*, int);
template <int DIM>
GatherMixedKernel pick_gather_mixed(int K) {
    switch (slot_width(K)) {
        case 4: return k_gather_mixed<DIM, 4>;
        case 6: return k_gather_mixed<DIM, 6>;
        default: return k_gather_mixed<DIM, 8>;
    }
}
template <int DIM, bool RUS>
TileMixedKernel pick_tile_mixed1(bool recon, bool visc, bool species, int K) {
    if (recon)
        return visc ? pick_tile_mixed2<DIM, true, true, RUS>(species, K) : pick_tile_mixed2<DIM, true, false, RUS>(species, K);
    return visc ? pick_tile_mixed2<DIM, false, true, RUS>(species, K) : pick_tile_mixed2<DIM, false, false, RUS>(species, K);
}
template <int DIM>
TileMixedKernel pick_tile_mixed(bool recon, bool visc, bool species, int K, bool rus) {
    return rus ? pick_tile_mixed1<DIM, true>(recon, visc, species, K) : pick_tile_mixed1<DIM, false>(recon, visc, species, K);
}
using VjpKernel = void (*)(MeshView<double>, Phys, const double*, const double*, const int32_t*, const int32_t*,
                          double*);
using VjpThetaKernel = void (*)(MeshView<double>, Phys, const double*, const double*, int64_t, const int32_t*,
                               const int32_t*, const int32_t*, const double*, const double*, double*);
template <int DIM>
VjpKernel pick_vjp(bool rusanov, int K) {
    switch (slot_width(K)) {
        case 4: return rusanov ? k_vjp<DIM, true, 4> : k_vjp<DIM, false, 4>;
        case 6: return rusanov ? k_vjp<DIM, true, 6> : k_vjp<DIM, false, 6>;
        default: return rusanov ? k_vjp<DIM, true, 8> : k_vjp<DIM, false, 8>;
    }
}
using VjpFacesKernel = void (*)(MeshView<double>, Phys, const double*, const double*, double*);
using VjpGatherKernel = void (*)(MeshView<double>, Phys, const double*, const double*, const int32_t*,
                                 const int32_t*, double*);
template <int DIM>
VjpFacesKernel pick_vjp_faces(bool rusanov) {
    return rusanov ? k_vjp_faces<DIM, true> : k_vjp_faces<DIM, false>;
}
template <int DIM>
VjpGatherKernel pick_vjp_gather(int K) {
    switch (slot_width(K)) {
        case 4: return k_vjp_gather<DIM, 4>;
        case 6: return k_vjp_gather<DIM, 6>;
        default: return k_vjp_gather<DIM, 8>;
    }
}
template <int DIM>
VjpThetaKernel pick_vjp_theta(bool rusanov) {
    return rusanov ? k_vjp_theta<DIM, true> : k_vjp_theta<DIM, false>;
}
using TileKernel = void (*)(MeshView<double>, Phys, const double*, const double*, const double*, const double*,
                           const int32_t*, const int32_t*, const int32_t*, int, double*, int);
template <int DIM, bool RECON, bool VISC, int KM>
TileKernel pick_tile3(bool species) {
    return species ? k_tile_flux<DIM, RECON, VISC, true, KM> : k_tile_flux<DIM, RECON, VISC, false, KM>;
}
template <int DIM, bool RECON, bool VISC>
TileKernel pick_tile2(bool species, int K) {
    switch (slot_width(K)) {
        case 4: return pick_tile3<DIM, RECON, VISC, 4>(species);
        case 6: return pick_tile3<DIM, RECON, VISC, 6>(species);
        default: return pick_tile3<DIM, RECON, VISC, 8>(species);
    }
}
template <int DIM>
TileKernel pick_tile(bool recon, bool visc, bool species, int K) {
    if (recon) return visc ? pick_tile2<DIM, true, true>(species, K) : pick_tile2<DIM, true, false>(species, K);
    return visc ? pick_tile2<DIM, false, true>(species, K) : pick_tile2<DIM, false, false>(species, K);
}
template <int DIM>
GatherKernel pick_gather(bool species, int K) {
    switch (slot_width(K)) {
        case 4: return species ? k_gather<DIM, true, 4> : k_gather<DIM, false, 4>;
        case 6: return species ? k_gather<DIM, true, 6> : k_gather<DIM, false, 6>;
        default: return species ? k_gather<DIM, true, 8> : k_gather<DIM, false, 8>;
    }
}

using FusedKernel = void (*)(MeshView<double>, Phys, const double*, FusedView, double*, int);
template <int DIM, int KM>
FusedKernel pick_fused_k(bool recon, bool visc, int lim) {
    if (lim == kLimPetsc) return nullptr;
    if (recon) {
        if (visc) return lim == kLimBJ ? k_fused<DIM, true, true, kLimBJ, KM> : k_fused<DIM, true, true, kLimNone, KM>;
        return lim == kLimBJ ? k_fused<DIM, true, false, kLimBJ, KM> : k_fused<DIM, true, false, kLimNone, KM>;
    }
    return visc ? k_fused<DIM, false, true, kLimNone, KM> : nullptr;
}
// Fused kernels exist for quads / triangles (2-D, 4 slots), tetrahedra (4) and hexahedra, prisms, pyramids (6).
template <int DIM>
FusedKernel pick_fused(bool recon, bool visc, int lim, int K) {
    const int KM = slot_width(K);
    if (KM == 4) return pick_fused_k<DIM, 4>(recon, visc, lim);
    if constexpr (DIM == 3)
        if (KM == 6) return pick_fused_k<DIM, 6>(recon, visc, lim);
    return nullptr;
}

using MarchKernel = void (*)(MeshView<double>, Phys, const double*, MarchView, double*, int);
template <int DIM, int KM>
MarchKernel pick_march_k(bool recon, bool visc, int lim) {
    if (lim == kLimPetsc) return nullptr;
    if (recon) {
        if (visc) return lim == kLimBJ ? k_march<DIM, true, true, kLimBJ, KM> : k_march<DIM, true, true, kLimNone, KM>;
        return lim == kLimBJ ? k_march<DIM, true, false, kLimBJ, KM> : k_march<DIM, true, false, kLimNone, KM>;
    }
    return visc ? k_march<DIM, false, true, kLimNone, KM> : nullptr;
}
// Marching kernels exist where the fused ones do.
template <int DIM>
MarchKernel pick_march(bool recon, bool visc, int lim, int K) {
    const int KM = slot_width(K);
    if (KM == 4) return pick_march_k<DIM, 4>(recon, visc, lim);
    if constexpr (DIM == 3)
        if (KM == 6) return pick_march_k<DIM, 6>(recon, visc, lim);
    return nullptr;
}

__global__ __launch_bounds__(kBlock) void k_zero_rows(double* __restrict__ R, int64_t N, int64_t start, int nc) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t cnt = N - start;
    if (i >= cnt * nc) return;
    const int64_t p = i / cnt, c = start + i % cnt;
    R[p * N + c] = 0.0;
}

__global__ __launch_bounds__(kBlock) void k_pack(const double* __restrict__ src, double* __restrict__ dst, int64_t N, int nc,
                                                 int to_soa) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N * nc) return;
    if (to_soa) {  // i indexes the SoA destination (coalesced writes)
        const int64_t p = i / N, c = i % N;
        dst[i] = src[c * nc + p];
    } else {  // i indexes the AoS destination
        const int64_t c = i / nc, p = i % nc;
        dst[i] = src[p * N + c];
    }
}

__global__ __launch_bounds__(kBlock) void k_check_state(const double* __restrict__ U, int64_t N, int64_t NI, int nc,
                                                        unsigned long long* __restrict__ bad) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int is_bad = 0;
    if (c < NI) {
        for (int p = 0; p < nc; p++)
            if (!isfinite(U[p * N + c])) is_bad = 1;
        if (!(U[c] > 0.0)) is_bad = 1;
    }
    // wavefront ballot: one integer atomic per wave (order-independent count)
    const unsigned long long mask = __ballot(is_bad);
    if ((threadIdx.x & 63) == 0 && mask) atomicAdd(bad, (unsigned long long)__popcll(mask));
}

// NavierStokesTransport::ComputeCflTimeStep per interior cell (navierStokesTransport.cpp:246-281, pgsAlpha = 1):
// dt = cfl * 2 r_min / (a + sum_d |rhoU_d| / rho), a through the PerfectGas T route (perfectGas.cpp:76-86,165-174),
// reduced to one minimum per block (min is order-free, so the result does not depend on the launch shape; NaN
// cells are skipped).
__device__ inline double block_min(double v) {
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o));
    __shared__ double wmin[kBlock / 64];
    if ((threadIdx.x & 63) == 0) wmin[threadIdx.x >> 6] = v;
    __syncthreads();
    double m = wmin[0];
    for (int i = 1; i < (int)(blockDim.x >> 6); i++) m = fmin(m, wmin[i]);
    return m;
}
__global__ __launch_bounds__(kBlock) void k_cfl_partial(const double* __restrict__ U, const double* __restrict__ rmin,
                                                        int64_t N, int64_t NI, int dim, Phys ph, double cfl,
                                                        double alpha, double* __restrict__ partial) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double dt = 1e10;  // utilities::Constants::large, the reference's starting minimum
    if (c < NI) {
        const double rho = U[c];
        double ke = 0.0, velSum = 0.0;
        for (int d = 0; d < dim; d++) {
            const double v = U[(kRhoU + d) * N + c] / rho;
            ke += v * v;
            velSum += fabs(U[(kRhoU + d) * N + c]) / rho;
        }
        ke *= 0.5;
        const double T = (U[kRhoE * N + c] / rho - ke) / ph.cv;
        const double p = ph.R * rho * T;
        const double a = sqrt(ph.gamma * p / rho);
        dt = cfl * (2.0 * rmin[c]) / (a / alpha + velSum);
    }
    const double m = block_min(dt);
    if (threadIdx.x == 0) partial[blockIdx.x] = m;
}
__global__ __launch_bounds__(kBlock) void k_min_final(const double* __restrict__ partial, int64_t n,
                                                      double* __restrict__ out) {
    double v = 1e10;
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) v = fmin(v, partial[i]);
    const double m = block_min(v);
    if (threadIdx.x == 0) out[0] = m;
}
// NavierStokesTransport::ComputeConductionTimeStep and ComputeViscousDiffusionTimeStep
// (navierStokesTransport.cpp:287-450): per interior cell |f (2 r_min)^2 / alpha| with alpha = k / (rho cv) and
// |f (2 r_min)^2 / nu| with nu = mu / rho, k and mu of the transport model at the cell temperature (the aux T);
// block minima of both into partial[0, nb) and partial[nb, 2 nb).
__global__ __launch_bounds__(kBlock) void k_diffusion_dt_partial(const double* __restrict__ U,
                                                                 const double* __restrict__ rmin, int64_t N,
                                                                 int64_t NI, int dim, Phys ph, double fcond,
                                                                 double fvisc, double* __restrict__ partial) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double dc = 1e10, dv = 1e10;  // utilities::Constants::large
    if (c < NI) {
        const double rho = U[c];
        double ke = 0.0;
        for (int d = 0; d < dim; d++) {
            const double v = U[(kRhoU + d) * N + c] / rho;
            ke += v * v;
        }
        ke *= 0.5;
        const double T = (U[kRhoE * N + c] / rho - ke) / ph.cv;
        double k = ph.k, mu = ph.mu;
        if (ph.suth) {  // ph.k, ph.mu are 1 or 0 per enabled property
            const double m = sutherland_mu(T);
            k *= m * ph.cp / kSuthPr;
            mu *= m;
        }
        const double dx2 = (2.0 * rmin[c]) * (2.0 * rmin[c]);
        const double alpha = k / (rho * ph.cv), nu = mu / rho;
        dc = fmin(dc, fabs(fcond * dx2 / alpha));
        dv = fmin(dv, fabs(fvisc * dx2 / nu));
    }
    const double mc = block_min(dc);
    const double mv = block_min(dv);
    if (threadIdx.x == 0) {
        partial[blockIdx.x] = mc;
        partial[gridDim.x + blockIdx.x] = mv;
    }
}
// PressureGradientScaling::UpdatePreconditioner reductions over the interior cells (pressureGradientScaling.cpp:55-104,
// PerfectGas T-route): r = {sum p, min p, max p, max a, max Mach}; a fixed reduction tree (deterministic sums),
// NaN entries skipped by the min / max.
constexpr int kPgsN = 5;
__device__ inline void block_reduce_pgs(double* r) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        r[0] += __shfl_xor(r[0], o);
        r[1] = fmin(r[1], __shfl_xor(r[1], o));
        r[2] = fmax(r[2], __shfl_xor(r[2], o));
        r[3] = fmax(r[3], __shfl_xor(r[3], o));
        r[4] = fmax(r[4], __shfl_xor(r[4], o));
    }
    __shared__ double wr[kPgsN][kBlock / 64];
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int i = 0; i < kPgsN; i++) wr[i][threadIdx.x >> 6] = r[i];
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kPgsN; i++) r[i] = wr[i][0];
    for (int w = 1; w < (int)(blockDim.x >> 6); w++) {
        r[0] += wr[0][w];
        r[1] = fmin(r[1], wr[1][w]);
        r[2] = fmax(r[2], wr[2][w]);
        r[3] = fmax(r[3], wr[3][w]);
        r[4] = fmax(r[4], wr[4][w]);
    }
}
__device__ inline void pgs_identity(double* r) {
    r[0] = 0.0;
    r[1] = DBL_MAX;
    r[2] = 0.0;
    r[3] = 0.0;
    r[4] = 0.0;
}
__global__ __launch_bounds__(kBlock) void k_pgs_partial(const double* __restrict__ U, int64_t N, int64_t NI, int dim,
                                                        Phys ph, double* __restrict__ partial) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double r[kPgsN];
    pgs_identity(r);
    if (c < NI) {
        const double rho = U[c];
        double vel2 = 0.0;
        for (int d = 0; d < dim; d++) {
            const double v = U[(kRhoU + d) * N + c] / rho;
            vel2 += v * v;
        }
        const double T = (U[kRhoE * N + c] / rho - 0.5 * vel2) / ph.cv;
        const double p = ph.R * rho * T;
        const double a = sqrt(ph.gamma * p / rho);
        r[0] = p;
        r[1] = p;
        r[2] = p;
        r[3] = a;
        r[4] = sqrt(vel2) / a;
    }
    block_reduce_pgs(r);
    if (threadIdx.x == 0)
#pragma unroll
        for (int i = 0; i < kPgsN; i++) partial[(int64_t)blockIdx.x * kPgsN + i] = r[i];
}
__global__ __launch_bounds__(kBlock) void k_pgs_final(const double* __restrict__ partial, int64_t n,
                                                      double* __restrict__ out) {
    double r[kPgsN];
    pgs_identity(r);
    for (int64_t b = threadIdx.x; b < n; b += blockDim.x) {
        const double* q = partial + b * kPgsN;
        r[0] += q[0];
        r[1] = fmin(r[1], q[1]);
        r[2] = fmax(r[2], q[2]);
        r[3] = fmax(r[3], q[3]);
        r[4] = fmax(r[4], q[4]);
    }
    block_reduce_pgs(r);
    if (threadIdx.x == 0)
#pragma unroll
        for (int i = 0; i < kPgsN; i++) out[i] = r[i];
}

// Forward-Euler stage update of the interior rows in FP64: U[p][c] += dt * R[p][c] (blockIdx.y = component).
__global__ __launch_bounds__(kBlock) void k_axpy(double dt, const double* __restrict__ R, double* __restrict__ U,
                                                 int64_t N, int64_t NI) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= NI) return;
    const int64_t i = (int64_t)blockIdx.y * N + c;
    U[i] += dt * R[i];
}

// ============================================================================================ host side

#define HIP_CHECK(x)                                                                                                 \
    do {                                                                                                             \
        hipError_t err_ = (x);                                                                                       \
        if (err_ != hipSuccess) throw std::runtime_error(std::string(#x) + ": " + hipGetErrorString(err_));          \
    } while (0)

template <typename T>
struct DevBuf {
    T* ptr = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() {
        if (ptr) (void)hipFree(ptr);
    }
    void alloc(size_t count) {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        n = count;
        if (count) HIP_CHECK(hipMalloc(&ptr, count * sizeof(T)));
    }
    void upload(const std::vector<T>& v) {
        alloc(v.size());
        if (!v.empty()) HIP_CHECK(hipMemcpy(ptr, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    }
    size_t bytes() const { return n * sizeof(T); }
};

// Slot-major copies of the per-side face geometry for the gradient kernels: for slot k (< KM, the compile-time
// slot width) of interior cell c, the neighbour across the face, the cell's least-squares weight vectors (cell
// and aux sets) and the limiter geometry (x_f - x_c for Barth-Jespersen, x_n - x_c for the PETSc family, both
// in G arithmetic from the G-rounded face records, as the face-major kernels form them).
template <typename G>
struct SlotGeometry {
    std::vector<int32_t> nbr;
    std::vector<G> w, wa, lim;
};
template <typename G>
void build_slot_geometry(const Layout& L, int KM, int lim, SlotGeometry<G>& out) {
    const int dim = L.dim;
    const int64_t NI = L.NI, F = L.F;
    out.nbr.assign((size_t)KM * NI, -1);
    if (!L.fw.empty()) out.w.assign((size_t)KM * dim * NI, G(0));
    if (!L.fwa.empty()) out.wa.assign((size_t)KM * dim * NI, G(0));
    if (lim != kLimNone) out.lim.assign((size_t)KM * dim * NI, G(0));
    for (int k = 0; k < L.K; k++)
        for (int64_t c = 0; c < NI; c++) {
            const int32_t e = L.ell[(size_t)k * NI + c];
            if (e < 0) continue;
            const int64_t f = e >> 1;
            const int s = e & 1;
            out.nbr[(size_t)k * NI + c] = L.fcell[(size_t)(1 - s) * F + f];
            for (int d = 0; d < dim; d++) {
                const size_t o = ((size_t)k * dim + d) * NI + c;
                const size_t a = (size_t)2 * dim * f + (size_t)s * dim + d;
                const size_t b = (size_t)2 * dim * f + (size_t)(1 - s) * dim + d;
                if (!out.w.empty()) out.w[o] = G(L.fw[a]);
                if (!out.wa.empty()) out.wa[o] = G(L.fwa[a]);
                if (lim == kLimBJ) out.lim[o] = G(L.fdx[a]);
                else if (lim == kLimPetsc) out.lim[o] = G(L.fdx[a]) - G(L.fdx[b]);
            }
        }
}

}  // namespace abl

struct abl_fv_solver {
    std::string last_error;
    abl_fv_physics_desc phys{};
    int form = 0;
    abl::Layout host;  // metadata only after upload (arrays released)
    abl::Phys ph{};
    abl::DevBuf<int32_t> ell, fcell, bface, bslot, slot_kind, snbr;
    abl::DevBuf<double> fna, fdx, fwa, inv_vol, gradC, gradA, slot_state;
    abl::DevBuf<double> sw, swa, slim;  // slot-major gradient geometry (FP64 form)
    abl::DevBuf<double> rec;            // cell records of the FP64 form (rec_pair blocks)
    // face fluxes (allocated when that strategy is selected): FP64 euler records [F][FS] and species pairs
    // [ns/2][F][2];
    // mixed form rows [2 (n_euler + n_species)][F]
    abl::DevBuf<double> face_flux_e, face_flux_s;
    abl::DevBuf<float> face_flux_m;
    abl::DevBuf<int32_t> tile_start, tile_faces, tile_slot;  // tile strategy (built when selected)
    int tile_T = 0, tile_nfm = 0;
    int64_t tile_n = 0;
    size_t tile_lds = 0;
    double tile_redundancy = 0.0;  // face evaluations of the tile kernel per face
    abl::TileKernel tile_kernel = nullptr;
    // fused strategy (built at creation when available)
    abl::DevBuf<int32_t> fu_start, fu_faces, fu_code, fu_slot, fu_hstart, fu_hcell, fu_hinfo;
    int fu_T = 0, fu_nfm = 0, fu_nem = 0;
    int64_t fu_n = 0;
    size_t fu_lds = 0;
    abl::FusedKernel fused_kernel = nullptr;
    std::string fused_unavailable;  // why the tiles were not built, when it failed
    // marching strategy (built at creation when available)
    abl::DevBuf<int32_t> mr_seq, mr_pro, mr_stile, mr_snext, mr_fstart, mr_ffaces, mr_fcode, mr_pfstart, mr_pffaces,
        mr_pfcode, mr_hstart, mr_hcell, mr_hinfo, mr_gslot;
    int mr_T = 0, mr_nfm = 0, mr_nem = 0;
    int64_t mr_nseq = 0;
    size_t mr_lds = 0;
    abl::MarchKernel march_kernel = nullptr;
    std::string march_unavailable;  // why the plan was not built, when it failed
    // mixed form: FP32 geometry and gradients
    abl::DevBuf<float> fna_f, fdx_f, fwa_f;
    abl::DevBuf<float> sw_f, swa_f, slim_f;  // slot-major gradient geometry (mixed form)
    abl::GradKernelT<float> grad_kernel_f = nullptr;
    abl::MixedFluxKernel flux_kernel_mixed = nullptr;
    abl::TileMixedKernel tile_kernel_mixed = nullptr;
    abl::FaceFluxMixedKernel face_flux_kernel_mixed = nullptr;
    abl::GatherMixedKernel gather_kernel_mixed = nullptr;
    abl::DevBuf<float> recf;            // cell records of the mixed form (recf_pair blocks)
    abl::DevBuf<float> gS_f, gY_f;      // mixed form species gradients (sgp_word pairs in gS_f; gY_f unused)
    // adjoint
    abl::DevBuf<int32_t> ghost_slot;
    abl::DevBuf<double> slot_param, slot_Y, theta_contrib;
    abl::DevBuf<double> theta_sum;    // [3 kMaxBcSlots] per-slot sums (k_theta_reduce)
    std::vector<double> h_slot_param, h_slot_Y;
    abl::VjpKernel vjp_kernel = nullptr;         // cell-centric first-order VJP (allocation-free entry)
    abl::VjpFacesKernel vjp_faces = nullptr;     // face-centric first-order VJP: face rows, then per-cell sums
    abl::VjpGatherKernel vjp_gather = nullptr;
    abl::DevBuf<double> vjp_rows;                // [2 NE + 2 + ns][F] face rows of vjp_faces (first call allocates)
    // second-order adjoint: kernels, mirror-slot table, workspace (allocated by the first such call)
    abl::AdjFacesKernel adj_faces = nullptr;
    abl::DevBuf<double> adjW, adjGL;  // face adjoint records, boundary-face species extras
    abl::AdjCellsKernel adj_cells_e = nullptr, adj_cells_s = nullptr;
    abl::AdjGatherKernel adj_gather = nullptr;
    abl::DevBuf<int32_t> mirror;
    abl::DevBuf<double> adjUghost, adjUself, adjContrib;
    abl::VjpThetaKernel vjp_theta_kernel = nullptr;
    abl::GradKernel grad_kernel = nullptr;
    abl::SpeciesGradKernelT<double> species_kernel = nullptr;
    abl::SpeciesGradKernelT<float> species_kernel_f = nullptr;
    abl::FluxKernel flux_kernel = nullptr;
    abl::FaceFluxKernel face_flux_kernel = nullptr;
    abl::GatherKernel gather_kernel = nullptr;
    int strategy = ABL_FV_STRATEGY_TILE;
    abl::DevBuf<unsigned long long> counter;
    abl::DevBuf<double> rmin, cfl_partial, cfl_out;  // MIN_CELL_RADIUS per interior cell; CFL reduction scratch
    abl::DevBuf<double> dt_partial;                  // conduction / viscous time-step block minima [2][blocks]
    // PressureGradientScaling attached by abl_fv_set_pgs: its state (alpha in ph.pgs_alpha) and reduction scratch
    bool pgs_on = false;
    abl_fv_pgs_desc pgs{};
    double pgs_max_mach = 0.0;
    abl::DevBuf<double> pgs_partial, pgs_out;
    std::vector<int32_t> h_slot_kind;        // ABL_FV_BC_* per slot (as set)
    std::vector<char> h_slot_at_face;        // INFLOW slots with enforceAtFace
    std::vector<int> bc_slots;               // slots that have boundary faces (set at creation)
    std::vector<double> h_slot_state;
    std::vector<std::array<double, 3>> h_slot_theta;
    int64_t n_bfaces = 0;
    bool recon = false, visc = false;
    // per-phase event timing
    bool profiling = false;
    struct PendingEvent {
        int phase;
        hipEvent_t start, stop;
    };
    std::vector<PendingEvent> pending;
    double phase_ms[ABL_FV_N_PHASES] = {0};
    int64_t phase_launches[ABL_FV_N_PHASES] = {0};

    // Launches `launch()` bracketed by events when profiling is on.
    template <typename Fn>
    void timed(int phase, hipStream_t st, Fn&& launch) {
        if (!profiling) {
            launch();
            return;
        }
        PendingEvent e{phase, nullptr, nullptr};
        HIP_CHECK(hipEventCreate(&e.start));
        HIP_CHECK(hipEventCreate(&e.stop));
        HIP_CHECK(hipEventRecord(e.start, st));
        launch();
        HIP_CHECK(hipEventRecord(e.stop, st));
        pending.push_back(e);
    }
    void drain_events() {
        for (auto& e : pending) {
            float ms = 0.f;
            HIP_CHECK(hipEventSynchronize(e.stop));
            HIP_CHECK(hipEventElapsedTime(&ms, e.start, e.stop));
            phase_ms[e.phase] += ms;
            phase_launches[e.phase] += 1;
            (void)hipEventDestroy(e.start);
            (void)hipEventDestroy(e.stop);
        }
        pending.clear();
    }
    ~abl_fv_solver() {
        for (auto& e : pending) {
            (void)hipEventSynchronize(e.stop);
            (void)hipEventDestroy(e.start);
            (void)hipEventDestroy(e.stop);
        }
    }

    abl::MeshView<double> view() const {
        abl::MeshView<double> v;
        v.N = host.N;
        v.NI = host.NI;
        v.F = host.F;
        v.K = host.K;
        v.ell = ell.ptr;
        v.fcell = fcell.ptr;
        v.fna = fna.ptr;
        v.fdx = fdx.ptr;
        v.fw = nullptr;
        v.fwa = fwa.ptr;
        v.inv_vol = inv_vol.ptr;
        v.snbr = snbr.ptr;
        v.sw = sw.ptr;
        v.swa = swa.ptr;
        v.slim = slim.ptr;
        return v;
    }
    abl::MeshView<float> view_f() const {
        abl::MeshView<float> v;
        v.N = host.N;
        v.NI = host.NI;
        v.F = host.F;
        v.K = host.K;
        v.ell = ell.ptr;
        v.fcell = fcell.ptr;
        v.fna = fna_f.ptr;
        v.fdx = fdx_f.ptr;
        v.fw = nullptr;
        v.fwa = fwa_f.ptr;
        v.inv_vol = inv_vol.ptr;
        v.snbr = snbr.ptr;
        v.sw = sw_f.ptr;
        v.swa = swa_f.ptr;
        v.slim = slim_f.ptr;
        return v;
    }
    // Face-flux storage of strategy 2 (see the members).
    void ensure_face_flux() {
        const int nrow = host.dim + 2 + host.ns;
        if (form == ABL_FV_FORM_FP64) {
            if (face_flux_e.ptr) return;
            face_flux_e.alloc((size_t)(((host.dim + 2) + 1) & ~1) * host.F);
            if (host.ns) {  // species pairs interleaved per face (k_gather); an odd count's last slot stays zero
                face_flux_s.alloc((size_t)(host.ns + 1) / 2 * 2 * host.F);
                HIP_CHECK(hipMemset(face_flux_s.ptr, 0, face_flux_s.bytes()));
            }
        } else {
            if (face_flux_m.ptr) return;
            face_flux_m.alloc((size_t)2 * nrow * host.F);
        }
    }
    // Tiles for strategy 3 (FP64 form): the largest tile (256, 128 or 64 cells) whose face fluxes fit the LDS
    // budget.  Needs the host cell -> face table, so it is built at creation.
    // only_T = 0: the largest of 256/128/64 cells within the default LDS budget; otherwise exactly only_T cells
    // within the workgroup LDS limit.  Returns false when no size fits (the strategy then stays unavailable).
    bool build_tiles(const abl::Layout& lay, int only_T = 0) {
        const int KM = abl::slot_width(lay.K);
        // LDS per face: FP64 euler + species fluxes, or the mixed form's two FP32 perturbation records
        const int per_face = form == ABL_FV_FORM_FP64 ? (lay.dim + 2 + lay.ns) * (int)sizeof(double)
                                                      : 2 * (lay.dim + 2 + lay.ns) * (int)sizeof(float);
        abl::Tiles tl;
        const size_t budget = only_T ? (size_t)abl::kTileLdsMax : (size_t)abl::kTileLdsDefault;
        for (int T = only_T ? only_T : abl::kTileDefault; T >= 64; T /= 2) {
            abl::build_tiles(lay, KM, T, tl);
            if ((size_t)tl.nf_max * per_face <= budget) break;
            tl.n = 0;
            if (only_T) break;
        }
        if (tl.n == 0) return false;
        if (form == ABL_FV_FORM_FP64 && tile_kernel)
            HIP_CHECK(hipFuncSetAttribute((const void*)tile_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                          (int)((size_t)tl.nf_max * per_face)));
        if (form != ABL_FV_FORM_FP64 && tile_kernel_mixed)
            HIP_CHECK(hipFuncSetAttribute((const void*)tile_kernel_mixed, hipFuncAttributeMaxDynamicSharedMemorySize,
                                          (int)((size_t)tl.nf_max * per_face)));
        tile_start.upload(tl.start);
        tile_faces.upload(tl.faces);
        tile_slot.upload(tl.slot);
        tile_T = tl.T;
        tile_n = tl.n;
        tile_nfm = tl.nf_max;
        tile_lds = (size_t)tl.nf_max * per_face;
        tile_redundancy = (double)tl.faces.size() / (double)std::max<int64_t>(host.F, 1);
        return true;
    }
    // Fused strategy tiles: the largest of 128 / 64 cells whose LDS (tile records, halo face entries, face fluxes)
    // leaves two workgroups per CU.  Needs the host cell -> face table and face offsets, so it is built at creation.
    // Returns false when the strategy is unavailable (no kernel for this handle, an ill-conditioned least-squares
    // stencil, or no tile size fits).
    bool build_fused(const abl::Layout& lay) {
        if (!fused_kernel) return false;
        if ((recon || visc) && !abl::lsq_stencils_full_rank(lay)) return false;
        const int KM = abl::slot_width(lay.K);
        const int ne = lay.dim + 2, nt = lay.dim * (lay.dim + 1) / 2 + lay.dim;
        const int rw = ne + (recon ? ne * lay.dim : 0) + (visc ? nt : 0);
        const int hw = ne + (visc ? lay.dim + nt : 0);
        abl::FusedTiles ft;
        size_t lds = 0;
        try {
            for (int T = 128; T >= 64; T /= 2) {
                abl::build_fused_tiles(lay, KM, T, ft);
                lds = sizeof(double) * ((size_t)rw * T + (size_t)hw * ft.nem + (size_t)ne * ft.nfm);
                if (lds <= (size_t)abl::kCuLdsBytes / 2) break;
                ft.n = 0;
            }
            if (ft.n == 0) return false;
            abl::check_fused_tiles(lay, KM, ft);
        } catch (const std::exception& e) {  // the strategy stays unavailable; the others are unaffected
            fused_unavailable = e.what();
            return false;
        }
        HIP_CHECK(hipFuncSetAttribute((const void*)fused_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        fu_start.upload(ft.fstart);
        fu_faces.upload(ft.ffaces);
        fu_code.upload(ft.fcode);
        fu_slot.upload(ft.fslot);
        fu_hstart.upload(ft.hstart);
        fu_hcell.upload(ft.hcell);
        fu_hinfo.upload(ft.hinfo);
        fu_T = ft.T;
        fu_n = ft.n;
        fu_nfm = ft.nfm;
        fu_nem = ft.nem;
        fu_lds = lds;
        return true;
    }
    // Marching plan: tiles of 64 cells, sequences sized so that the workgroups fill the device in about four rounds
    // at three workgroups per CU (LDS of ~51 KB per workgroup for the 3-D Navier-Stokes records).  Needs the host
    // cell -> face table and face offsets, so it is built at creation; false when unavailable.
    bool build_march(const abl::Layout& lay) {
        if (!march_kernel) return false;
        if ((recon || visc) && !abl::lsq_stencils_full_rank(lay)) return false;
        const int KM = abl::slot_width(lay.K);
        const int ne = lay.dim + 2, nt = lay.dim * (lay.dim + 1) / 2 + lay.dim;
        const int rw = ne + (recon ? ne * lay.dim : 0) + (visc ? nt : 0);
        const int hw = ne + (visc ? lay.dim + nt : 0);
        const int T = 64;
        int dev = 0, cus = 256;
        HIP_CHECK(hipGetDevice(&dev));
        HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        const int64_t tiles = (lay.NI + T - 1) / T;
        const int64_t target = std::max<int64_t>(1, (int64_t)cus * 3 * 4);
        const int max_steps = (int)std::max<int64_t>(2, (tiles + target - 1) / target);
        abl::MarchPlan mp;
        try {
            abl::build_march_plan(lay, KM, T, max_steps, mp);
            abl::check_march_plan(lay, KM, mp);
        } catch (const std::exception& e) {  // the strategy stays unavailable; the others are unaffected
            march_unavailable = e.what();
            return false;
        }
        const size_t lds = sizeof(double) * ((size_t)2 * rw * T + (size_t)hw * mp.nem + (size_t)2 * ne * mp.nfm);
        if (lds > (size_t)abl::kCuLdsBytes) return false;
        HIP_CHECK(hipFuncSetAttribute((const void*)march_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        mr_seq.upload(mp.seq);
        mr_pro.upload(mp.pro);
        mr_stile.upload(mp.stile);
        mr_snext.upload(mp.snext);
        mr_fstart.upload(mp.fstart);
        mr_ffaces.upload(mp.ffaces);
        mr_fcode.upload(mp.fcode);
        mr_pfstart.upload(mp.pfstart);
        mr_pffaces.upload(mp.pffaces);
        mr_pfcode.upload(mp.pfcode);
        mr_hstart.upload(mp.hstart);
        mr_hcell.upload(mp.hcell);
        mr_hinfo.upload(mp.hinfo);
        mr_gslot.upload(mp.gslot);
        mr_T = mp.T;
        mr_nfm = std::max(mp.nfm, 1);
        mr_nem = std::max(mp.nem, 1);
        mr_nseq = mp.n_seq;
        mr_lds = sizeof(double) * ((size_t)2 * rw * T + (size_t)hw * mr_nem + (size_t)2 * ne * mr_nfm);
        return true;
    }
    abl::MarchView march_view() const {
        abl::MarchView v;
        v.seq = mr_seq.ptr;
        v.pro = mr_pro.ptr;
        v.stile = mr_stile.ptr;
        v.snext = mr_snext.ptr;
        v.fstart = mr_fstart.ptr;
        v.ffaces = mr_ffaces.ptr;
        v.fcode = mr_fcode.ptr;
        v.pfstart = mr_pfstart.ptr;
        v.pffaces = mr_pffaces.ptr;
        v.pfcode = mr_pfcode.ptr;
        v.hstart = mr_hstart.ptr;
        v.hcell = mr_hcell.ptr;
        v.hinfo = mr_hinfo.ptr;
        v.gslot = mr_gslot.ptr;
        v.T = mr_T;
        v.nfm = mr_nfm;
        v.nem = mr_nem;
        return v;
    }
    abl::FusedView fused_view() const {
        abl::FusedView v;
        v.fstart = fu_start.ptr;
        v.ffaces = fu_faces.ptr;
        v.fcode = fu_code.ptr;
        v.fslot = fu_slot.ptr;
        v.hstart = fu_hstart.ptr;
        v.hcell = fu_hcell.ptr;
        v.hinfo = fu_hinfo.ptr;
        v.T = fu_T;
        v.nfm = fu_nfm;
        v.nem = fu_nem;
        return v;
    }
    // Default strategy.  The tile kernel evaluates faces on tile borders twice and holds its face fluxes in LDS;
    // the face-flux strategy evaluates every face once but writes and re-reads the fluxes through HBM.  Tiles win
    // while the border overhead stays below ~1.8 evaluations per face (rows of a structured box: 1.2-1.7) and the
    // LDS footprint leaves at least two waves per SIMD for latency hiding; cell orders with little locality within
    // a tile (bandwidth-reducing orders of unstructured meshes: ~2 evaluations per face) and wide species LDS rows
    // go to face flux.
    // The fused strategy is offered on request only: on the 256^3 hex workload it took 8.3 ms per RHS against
    // 5.3 ms for gradients + tiles (two workgroups per CU by LDS and registers; its phases wait on memory).
    int auto_strategy() const {
        if (tile_n > 0) {
            const double wgs_per_cu = std::floor((double)abl::kCuLdsBytes / (double)std::max<size_t>(tile_lds, 1));
            const double waves_per_simd = wgs_per_cu * (tile_T / 64.0) / 4.0;
            if (tile_redundancy <= 1.8 && waves_per_simd >= 2.0) return ABL_FV_STRATEGY_TILE;
        }
        return ABL_FV_STRATEGY_FACE_FLUX;
    }
    int64_t device_bytes() const {
        return (int64_t)(ell.bytes() + fcell.bytes() + bface.bytes() + bslot.bytes() + slot_kind.bytes() + fna.bytes() + fdx.bytes() + face_flux_e.bytes() + face_flux_s.bytes() + face_flux_m.bytes() + tile_start.bytes() + tile_faces.bytes() + tile_slot.bytes() + fu_start.bytes() + fu_faces.bytes() + fu_code.bytes() + fu_slot.bytes() + fu_hstart.bytes() + fu_hcell.bytes() + fu_hinfo.bytes() + mr_seq.bytes() + mr_pro.bytes() + mr_stile.bytes() + mr_snext.bytes() + mr_fstart.bytes() + mr_ffaces.bytes() + mr_fcode.bytes() + mr_pfstart.bytes() + mr_pffaces.bytes() + mr_pfcode.bytes() + mr_hstart.bytes() + mr_hcell.bytes() + mr_hinfo.bytes() + mr_gslot.bytes() + snbr.bytes() + vjp_rows.bytes() + sw.bytes() + swa.bytes() + slim.bytes() + sw_f.bytes() + swa_f.bytes() + slim_f.bytes() + fwa.bytes() + inv_vol.bytes() + gradC.bytes() +
                         gradA.bytes() + rec.bytes() + adjUghost.bytes() + adjUself.bytes() + adjContrib.bytes() + adjW.bytes() + adjGL.bytes() + mirror.bytes() + recf.bytes() + gS_f.bytes() + gY_f.bytes() + rmin.bytes() + cfl_partial.bytes() + dt_partial.bytes() + cfl_out.bytes() + slot_state.bytes() + counter.bytes() + fna_f.bytes() + fdx_f.bytes() +
                         fwa_f.bytes());
    }
};

namespace {

int fail(abl_fv_handle h, const std::string& msg) {
    if (h) h->last_error = msg;
    abl::set_global_error(msg);
    return 1;
}

inline unsigned grid_for(int64_t n) { return (unsigned)((n + abl::kBlock - 1) / abl::kBlock); }

// Every slot that has boundary faces needs a condition: an unset slot would leave its ghost cells undefined.
// Returns an empty string when all are set.
std::string unset_boundary(abl_fv_handle h) {
    for (int v : h->bc_slots)
        if (h->h_slot_kind[v] == ABL_FV_BC_NONE)
            return "boundary slot " + std::to_string(v) +
                   " has boundary faces but no boundary condition (abl_fv_set_boundary)";
    return std::string();
}

// theta_bar is the adjoint of the inflow states: with no inflow slot every contribution is zero and the theta_bar
// passes (and their copy back to the host) are skipped.
bool has_inflow(abl_fv_handle h) {
    for (int v : h->bc_slots)
        if (h->h_slot_kind[v] == ABL_FV_BC_INFLOW) return true;
    return false;
}

}  // namespace

extern "C" const char* abl_fv_last_error(abl_fv_handle h) {
    if (h) return h->last_error.c_str();
    return abl::g_last_error.c_str();
}

extern "C" int abl_fv_create(const abl_fv_mesh_desc* mesh, const abl_fv_physics_desc* physics, int32_t form,
                             abl_fv_handle* out) {
    if (!mesh || !physics || !out) return fail(nullptr, "abl_fv_create: null argument");
    *out = nullptr;
    if (form != ABL_FV_FORM_FP64 && form != ABL_FV_FORM_MIXED) return fail(nullptr, "abl_fv_create: unknown form");
    const abl_fv_physics_desc& p = *physics;
    if (!(p.gamma > 1.0) || !(p.Rgas > 0.0)) return fail(nullptr, "abl_fv_create: PerfectGas needs gamma > 1 and Rgas > 0");
    if (p.flux_calculator < ABL_FV_FLUX_NONE || p.flux_calculator > ABL_FV_FLUX_AUSMPUP)
        return fail(nullptr, "abl_fv_create: unknown flux calculator");
    if (p.reconstruction != ABL_FV_RECON_NONE && p.reconstruction != ABL_FV_RECON_LEASTSQUARES)
        return fail(nullptr, "abl_fv_create: unknown reconstruction");
    if (p.limiter < ABL_FV_LIMITER_NONE || p.limiter > ABL_FV_LIMITER_BARTH_JESPERSEN)
        return fail(nullptr, "abl_fv_create: unknown limiter");
    if (form == ABL_FV_FORM_MIXED && p.flux_calculator == ABL_FV_FLUX_AUSMPUP)
        return fail(nullptr, "abl_fv_create: the mixed form has no AusmpUp flux calculator (create the FP64 form)");
    if (p.transport != ABL_FV_TRANSPORT_CONSTANT && p.transport != ABL_FV_TRANSPORT_SUTHERLAND)
        return fail(nullptr, "abl_fv_create: unknown transport model");
    const bool suth = p.transport == ABL_FV_TRANSPORT_SUTHERLAND;
    if (suth && (p.transport_disabled & ~(ABL_FV_TRANSPORT_CONDUCTIVITY | ABL_FV_TRANSPORT_VISCOSITY |
                                         ABL_FV_TRANSPORT_DIFFUSIVITY)) != 0)
        return fail(nullptr, "abl_fv_create: transport_disabled holds bits other than ABL_FV_TRANSPORT_*");
    if (suth && form == ABL_FV_FORM_MIXED)
        return fail(nullptr, "abl_fv_create: the mixed form has no Sutherland transport (create the FP64 form)");
    std::unique_ptr<abl_fv_solver> s(new abl_fv_solver());
    try {
        s->phys = p;
        s->form = form;
        abl::build_layout(*mesh, p, s->host);
        abl::Layout& L = s->host;
        s->recon = p.reconstruction == ABL_FV_RECON_LEASTSQUARES && p.flux_calculator != ABL_FV_FLUX_NONE;
        abl::Phys& ph = s->ph;
        ph.gamma = p.gamma;
        ph.R = p.Rgas;
        ph.cv = p.Rgas / (p.gamma - 1.0);
        ph.icv = 1.0 / ph.cv;
        ph.cp = p.gamma * p.Rgas / (p.gamma - 1.0);
        ph.suth = suth;
        if (suth) {  // record scales 1 or 0 per enabled property; the faces apply mu(T_face)
            ph.mu = (p.transport_disabled & ABL_FV_TRANSPORT_VISCOSITY) ? 0.0 : 1.0;
            ph.k = (p.transport_disabled & ABL_FV_TRANSPORT_CONDUCTIVITY) ? 0.0 : 1.0;
            ph.diff = (p.transport_disabled & ABL_FV_TRANSPORT_DIFFUSIVITY) ? 0.0 : 1.0;
        } else {
            ph.mu = p.mu;
            ph.k = p.k;
            ph.diff = p.diff;
        }
        s->visc = ph.k != 0.0 || ph.mu != 0.0 || ph.diff != 0.0;
        ph.max_lim_grad = p.max_lim_grad;
        ph.flux = p.flux_calculator;
        ph.minf = p.ausmpup_minf;
        ph.compact = p.compact_species != 0;
        ph.pgs_alpha = 1.0;
        ph.recon = s->recon;
        ph.limiter = p.limiter;
        ph.ns = p.n_species;
        ph.visc = s->visc;
        ph.sdiff = s->visc && p.n_species > 0;
        const int lim = p.limiter == ABL_FV_LIMITER_NONE ? abl::kLimNone
                        : p.limiter == ABL_FV_LIMITER_BARTH_JESPERSEN ? abl::kLimBJ : abl::kLimPetsc;
        const bool species = p.n_species > 0;
        if (mesh->dim == 3) {
            s->species_kernel = abl::pick_species_gradients<double, 3>(L.K);
            s->species_kernel_f = abl::pick_species_gradients<float, 3>(L.K);
            s->grad_kernel = abl::pick_gradients<double, 3>(s->recon, s->visc, lim, L.K);
            s->flux_kernel = abl::pick_flux<3>(s->recon, s->visc, species);
            s->face_flux_kernel = abl::pick_face_flux<3>(s->recon, s->visc, species);
            s->tile_kernel = abl::pick_tile<3>(s->recon, s->visc, species, L.K);
            s->gather_kernel = abl::pick_gather<3>(species, L.K);
            if (form == ABL_FV_FORM_FP64 && !species && !suth) s->fused_kernel = abl::pick_fused<3>(s->recon, s->visc, lim, L.K);
            if (form == ABL_FV_FORM_FP64 && !species && !suth) s->march_kernel = abl::pick_march<3>(s->recon, s->visc, lim, L.K);
        } else {
            s->species_kernel = abl::pick_species_gradients<double, 2>(L.K);
            s->species_kernel_f = abl::pick_species_gradients<float, 2>(L.K);
            s->grad_kernel = abl::pick_gradients<double, 2>(s->recon, s->visc, lim, L.K);
            s->flux_kernel = abl::pick_flux<2>(s->recon, s->visc, species);
            s->face_flux_kernel = abl::pick_face_flux<2>(s->recon, s->visc, species);
            s->tile_kernel = abl::pick_tile<2>(s->recon, s->visc, species, L.K);
            s->gather_kernel = abl::pick_gather<2>(species, L.K);
            if (form == ABL_FV_FORM_FP64 && !species && !suth) s->fused_kernel = abl::pick_fused<2>(s->recon, s->visc, lim, L.K);
            if (form == ABL_FV_FORM_FP64 && !species && !suth) s->march_kernel = abl::pick_march<2>(s->recon, s->visc, lim, L.K);
        }

        s->ell.upload(L.ell);
        s->fcell.upload(L.fcell);
        s->inv_vol.upload(L.inv_vol);
        const int64_t N = L.N;
        const int nt = L.dim * (L.dim + 1) / 2 + L.dim;
        if (form == ABL_FV_FORM_FP64) {
            s->fna.upload(L.fna);
            s->fdx.upload(L.fdx);
            abl::SlotGeometry<double> sg;
            abl::build_slot_geometry<double>(L, abl::slot_width(L.K), lim, sg);
            s->snbr.upload(sg.nbr);
            s->sw.upload(sg.w);
            s->swa.upload(sg.wa);
            s->slim.upload(sg.lim);
            s->fwa.upload(L.fwa);
            if (s->recon || s->visc)
                s->rec.alloc((size_t)(L.dim == 3 ? abl::CellRec<3>::RS : abl::CellRec<2>::RS) * ((L.NI + 63) / 64 * 64));
            // species gradients: limited densityYi and Yi aux together, DIM word pairs per species and cell
            // (sgp_word); gradA stays unallocated in the FP64 form
            if ((s->recon || s->visc) && L.ns) {
                s->gradC.alloc((size_t)L.ns * 2 * L.dim * N);
                HIP_CHECK(hipMemset(s->gradC.ptr, 0, s->gradC.bytes()));  // ghost cells stay zero
            }
        } else {
            auto to_f = [](const std::vector<double>& v) { return std::vector<float>(v.begin(), v.end()); };
            s->fna_f.upload(to_f(L.fna));
            s->fdx_f.upload(to_f(L.fdx));
            abl::SlotGeometry<float> sg;
            abl::build_slot_geometry<float>(L, abl::slot_width(L.K), lim, sg);
            s->snbr.upload(sg.nbr);
            s->sw_f.upload(sg.w);
            s->swa_f.upload(sg.wa);
            s->slim_f.upload(sg.lim);
            s->fwa_f.upload(to_f(L.fwa));
            if (s->recon || s->visc)
                s->recf.alloc((size_t)(L.dim == 3 ? abl::CellRecF<3>::RS : abl::CellRecF<2>::RS) * ((L.NI + 63) / 64 * 64));
            // species gradients: limited densityYi and Yi aux together, DIM FP32 pairs per species and cell
            // (sgp_word); gY_f stays unallocated
            if ((s->recon || s->visc) && L.ns) {
                s->gS_f.alloc((size_t)L.ns * 2 * L.dim * N);
                HIP_CHECK(hipMemset(s->gS_f.ptr, 0, s->gS_f.bytes()));
            }
            const bool rus_m = p.flux_calculator == ABL_FV_FLUX_RUSANOV;
            if (mesh->dim == 3) {
                s->grad_kernel_f = abl::pick_gradients<float, 3>(s->recon, s->visc, lim, L.K);
                s->flux_kernel_mixed = abl::pick_flux_mixed<3>(s->recon, s->visc, rus_m);
                s->tile_kernel_mixed = abl::pick_tile_mixed<3>(s->recon, s->visc, species, L.K, rus_m);
                s->face_flux_kernel_mixed = abl::pick_face_flux_mixed<3>(s->recon, s->visc, species, rus_m);
                s->gather_kernel_mixed = abl::pick_gather_mixed<3>(L.K);
            } else {
                s->grad_kernel_f = abl::pick_gradients<float, 2>(s->recon, s->visc, lim, L.K);
                s->flux_kernel_mixed = abl::pick_flux_mixed<2>(s->recon, s->visc, rus_m);
                s->tile_kernel_mixed = abl::pick_tile_mixed<2>(s->recon, s->visc, species, L.K, rus_m);
                s->face_flux_kernel_mixed = abl::pick_face_flux_mixed<2>(s->recon, s->visc, species, rus_m);
                s->gather_kernel_mixed = abl::pick_gather_mixed<2>(L.K);
            }
        }
        std::vector<int32_t> bf, bs;
        {
            std::vector<char> seen(abl::kMaxBcSlots, 0);
            for (auto& b : L.bfaces) {
                bf.push_back(b.face);
                bs.push_back(b.slot);
                seen[b.slot] = 1;
            }
            for (int v = 0; v < abl::kMaxBcSlots; v++)
                if (seen[v]) s->bc_slots.push_back(v);
        }
        s->n_bfaces = (int64_t)bf.size();
        {
            std::vector<int32_t> gs(std::max<int64_t>(L.N - L.NI, 1), 0);
            for (auto& b : L.bfaces) gs[L.fcell[L.F + b.face] - L.NI] = b.slot;
            s->ghost_slot.upload(gs);
        }
        s->h_slot_param.assign((size_t)abl::kMaxBcSlots * 6, 0.0);
        s->h_slot_Y.assign((size_t)abl::kMaxBcSlots * std::max(L.ns, 1), 0.0);
        s->slot_param.upload(s->h_slot_param);
        s->slot_Y.upload(s->h_slot_Y);
        if (form == ABL_FV_FORM_FP64) {
            const bool rus = p.flux_calculator == ABL_FV_FLUX_RUSANOV;
            if (mesh->dim == 3) {
                s->vjp_kernel = abl::pick_vjp<3>(rus, L.K);
                s->vjp_faces = abl::pick_vjp_faces<3>(rus);
                s->vjp_gather = abl::pick_vjp_gather<3>(L.K);
                s->vjp_theta_kernel = abl::pick_vjp_theta<3>(rus);
                abl::pick_adj<3>(rus, L.K, s->adj_faces, s->adj_cells_e, s->adj_cells_s, s->adj_gather);
            } else {
                s->vjp_kernel = abl::pick_vjp<2>(rus, L.K);
                s->vjp_faces = abl::pick_vjp_faces<2>(rus);
                s->vjp_gather = abl::pick_vjp_gather<2>(L.K);
                s->vjp_theta_kernel = abl::pick_vjp_theta<2>(rus);
                abl::pick_adj<2>(rus, L.K, s->adj_faces, s->adj_cells_e, s->adj_cells_s, s->adj_gather);
            }
            if (s->recon || s->visc) {
                // mirror slots: for slot k of cell c, the slot of the same face in the neighbour's list
                const int KM = abl::slot_width(L.K);
                std::vector<int32_t> mir((size_t)KM * L.NI, -1), at0((size_t)L.F, -1), at1((size_t)L.F, -1);
                for (int k = 0; k < L.K; k++)
                    for (int64_t c = 0; c < L.NI; c++) {
                        const int32_t e = L.ell[(size_t)k * L.NI + c];
                        if (e < 0) continue;
                        ((e & 1) ? at1 : at0)[(size_t)(e >> 1)] = k;
                    }
                for (int k = 0; k < L.K; k++)
                    for (int64_t c = 0; c < L.NI; c++) {
                        const int32_t e = L.ell[(size_t)k * L.NI + c];
                        if (e < 0) continue;
                        mir[(size_t)k * L.NI + c] = ((e & 1) ? at0 : at1)[(size_t)(e >> 1)];
                    }
                s->mirror.upload(mir);
            }
        }
        s->bface.upload(bf);
        s->bslot.upload(bs);
        s->h_slot_kind.assign(abl::kMaxBcSlots, ABL_FV_BC_NONE);
        s->h_slot_at_face.assign(abl::kMaxBcSlots, 0);
        s->h_slot_state.assign((size_t)abl::kMaxBcSlots * L.nc, 0.0);
        s->h_slot_theta.assign(abl::kMaxBcSlots, {0.0, 0.0, 0.0});
        s->slot_kind.upload(s->h_slot_kind);
        s->slot_state.upload(s->h_slot_state);
        s->counter.alloc(1);
        {
            // MIN_CELL_RADIUS (finiteVolumeSolver.cpp:185-215): min over the cell's faces of |x_cell - x_face|
            std::vector<double> r((size_t)L.NI, 0.0);
            for (int64_t c = 0; c < L.NI; c++) {
                bool first = true;
                for (int k = 0; k < L.K; k++) {
                    const int32_t e = L.ell[(size_t)k * L.NI + c];
                    if (e < 0) break;
                    const int64_t f = e >> 1;
                    const int sd = e & 1;
                    double mag = 0.0;
                    for (int d = 0; d < L.dim; d++) {
                        const double v = L.fdx[(size_t)2 * L.dim * f + (size_t)sd * L.dim + d];
                        mag += v * v;
                    }
                    const double rad = std::sqrt(mag);
                    r[(size_t)c] = first ? rad : std::min(r[(size_t)c], rad);
                    first = false;
                }
            }
            s->rmin.upload(r);
            s->cfl_partial.alloc((size_t)std::max<int64_t>(grid_for(L.NI), 1));
            s->dt_partial.alloc((size_t)2 * std::max<int64_t>(grid_for(L.NI), 1));
            s->cfl_out.alloc(3);
        }
        s->build_tiles(L);
        s->build_fused(L);
        s->build_march(L);
        s->strategy = s->auto_strategy();
        abl::resolve_noalloc_scratch();
        if (s->strategy == ABL_FV_STRATEGY_FACE_FLUX) s->ensure_face_flux();
        // release the host copies of the big arrays
        L.ell.clear(); L.ell.shrink_to_fit();
        L.fcell.clear(); L.fcell.shrink_to_fit();
        L.fna.clear(); L.fna.shrink_to_fit();
        L.fdx.clear(); L.fdx.shrink_to_fit();
        L.fw.clear(); L.fw.shrink_to_fit();
        L.fwa.clear(); L.fwa.shrink_to_fit();
        L.inv_vol.clear(); L.inv_vol.shrink_to_fit();
    } catch (const std::exception& e) {
        return fail(nullptr, std::string("abl_fv_create: ") + e.what());
    }
    *out = s.release();
    return 0;
}

extern "C" int abl_fv_destroy(abl_fv_handle h) {
    delete h;
    return 0;
}

extern "C" int abl_fv_get_info(abl_fv_handle h, abl_fv_info* info) {
    if (!h || !info) return fail(h, "abl_fv_get_info: null argument");
    const abl::Layout& L = h->host;
    info->dim = L.dim;
    info->n_comp = L.nc;
    info->max_faces_per_cell = L.K;
    info->form = h->form;
    info->n_cells = L.N;
    info->n_interior = L.NI;
    info->n_faces = L.F;
    info->device_bytes = h->device_bytes();
    // algorithmic bytes: state read + residual write, per face (support pair, normal, centroid, LSQ weights
    // when reconstructing), per cell (1/V, centroid when reconstructing, cell->face table)
    const double sg = h->form == ABL_FV_FORM_FP64 ? 8.0 : 4.0;
    const double faces_per_cell = (double)L.F / (double)L.NI;
    const double per_face = 8.0 + L.dim * sg + L.dim * sg + (h->recon ? 2.0 * L.dim * sg : 0.0);
    const double per_cell = sg + (h->recon ? L.dim * sg : 0.0) + 4.0 * L.K + 4.0;
    info->alg_bytes_per_cell = 2.0 * L.nc * 8.0 + faces_per_cell * per_face + per_cell;
    return 0;
}

extern "C" int abl_fv_set_boundary(abl_fv_handle h, int32_t slot, const abl_fv_bc_desc* bc) {
    if (!h || !bc) return fail(h, "abl_fv_set_boundary: null argument");
    if (slot < 0 || slot >= abl::kMaxBcSlots) return fail(h, "abl_fv_set_boundary: slot out of range");
    if (bc->kind != ABL_FV_BC_INFLOW && bc->kind != ABL_FV_BC_OUTFLOW && bc->kind != ABL_FV_BC_NONE &&
        bc->kind != ABL_FV_BC_GHOST)
        return fail(h, "abl_fv_set_boundary: unknown boundary kind");
    if (bc->enforce_at_face && bc->kind != ABL_FV_BC_INFLOW)
        return fail(h, "abl_fv_set_boundary: enforce_at_face applies to INFLOW slots only");
    const abl::Layout& L = h->host;
    const int nc = L.nc, dim = L.dim;
    double* st = &h->h_slot_state[(size_t)slot * nc];
    for (int p = 0; p < nc; p++) st[p] = 0.0;
    if (bc->kind == ABL_FV_BC_INFLOW) {
        if (!(bc->rho > 0.0) || !(bc->T > 0.0)) return fail(h, "abl_fv_set_boundary: inflow needs rho > 0 and T > 0");
        double ke = 0.0;
        for (int d = 0; d < dim; d++) ke += bc->vel[d] * bc->vel[d];
        st[abl::kRho] = bc->rho;
        st[abl::kRhoE] = bc->rho * (bc->T * h->ph.cv + 0.5 * ke);
        for (int d = 0; d < dim; d++) st[abl::kRhoU + d] = bc->rho * bc->vel[d];
        for (int sp = 0; sp < L.ns; sp++) st[dim + 2 + sp] = bc->rho * (bc->Yi ? bc->Yi[sp] : 0.0);
    }
    h->h_slot_kind[slot] = bc->kind;
    h->h_slot_at_face[slot] = bc->kind == ABL_FV_BC_INFLOW && bc->enforce_at_face != 0;
    std::vector<int32_t> dev_kind(h->h_slot_kind);  // device codes: enforceAtFace inflow has its own
    for (int v = 0; v < abl::kMaxBcSlots; v++)
        if (dev_kind[v] == ABL_FV_BC_INFLOW && h->h_slot_at_face[v]) dev_kind[v] = abl::kBcInflowAtFace;
    {
        double* sp = &h->h_slot_param[(size_t)slot * 6];
        for (int i = 0; i < 6; i++) sp[i] = 0.0;
        if (bc->kind == ABL_FV_BC_INFLOW) {
            double um = 0.0;
            for (int d = 0; d < dim; d++) um += bc->vel[d] * bc->vel[d];
            um = std::sqrt(um);
            sp[0] = bc->rho;
            sp[1] = um;
            sp[2] = bc->T;
            for (int d = 0; d < dim; d++) sp[3 + d] = um > 0.0 ? bc->vel[d] / um : 0.0;
            for (int q = 0; q < L.ns; q++) h->h_slot_Y[(size_t)slot * L.ns + q] = bc->Yi ? bc->Yi[q] : 0.0;
        }
    }
    try {
        HIP_CHECK(hipMemcpy(h->slot_param.ptr, h->h_slot_param.data(), h->slot_param.bytes(), hipMemcpyHostToDevice));
        HIP_CHECK(hipMemcpy(h->slot_Y.ptr, h->h_slot_Y.data(), h->slot_Y.bytes(), hipMemcpyHostToDevice));
        HIP_CHECK(hipMemcpy(h->slot_kind.ptr, dev_kind.data(), h->slot_kind.bytes(), hipMemcpyHostToDevice));
        HIP_CHECK(hipMemcpy(h->slot_state.ptr, h->h_slot_state.data(), h->slot_state.bytes(), hipMemcpyHostToDevice));
    } catch (const std::exception& e) {
        return fail(h, e.what());
    }
    return 0;
}

extern "C" int abl_fv_rhs(abl_fv_handle h, double time, double* U, double* R, int32_t accumulate, abl_fv_stream stream) {
    (void)time;  // constant-state conditions do not depend on it; caller-filled ghosts were evaluated at it
    if (!h || !U || !R) return fail(h, "abl_fv_rhs: null argument");
    {
        const std::string bad = unset_boundary(h);
        if (!bad.empty()) return fail(h, "abl_fv_rhs: " + bad);
    }
    try {
        hipStream_t st = (hipStream_t)stream;
        const abl::Layout& L = h->host;
        if (h->n_bfaces) {
            h->timed(ABL_FV_PHASE_GHOST, st, [&] {
                hipLaunchKernelGGL(abl::k_ghost_fill, dim3(grid_for(h->n_bfaces)), dim3(abl::kBlock), 0, st, h->n_bfaces,
                                   h->bface.ptr, h->bslot.ptr, h->fcell.ptr, L.F, L.N, L.nc, h->slot_kind.ptr,
                                   h->slot_state.ptr, U);
            });
        }
        const dim3 grid(grid_for(L.NI)), block(abl::kBlock);
        if (h->form == ABL_FV_FORM_MIXED) {
            const auto vf = h->view_f();
            if (h->recon || h->visc)
                h->timed(ABL_FV_PHASE_GRADIENTS, st, [&] {
                    hipLaunchKernelGGL(h->grad_kernel_f, grid, block, 0, st, vf, h->ph, U, nullptr, nullptr,
                                       (void*)h->recf.ptr);
                });
            if (L.ns && (h->recon || h->visc))
                h->timed(ABL_FV_PHASE_SPECIES_GRADIENTS, st, [&] {
                    hipLaunchKernelGGL(h->species_kernel_f, grid, block, 0, st, vf, h->ph, U, h->gS_f.ptr, h->gY_f.ptr);
                });
            if (h->strategy == ABL_FV_STRATEGY_FACE_FLUX) {
                const int nrow = L.dim + 2 + L.ns;
                h->timed(ABL_FV_PHASE_FLUX, st, [&] {
                    hipLaunchKernelGGL(h->face_flux_kernel_mixed, dim3(grid_for(L.F)), block, 0, st, vf, h->ph, U,
                                       h->recf.ptr,
                                       h->gS_f.ptr, h->gY_f.ptr, h->face_flux_m.ptr, nrow);
                });
                h->timed(ABL_FV_PHASE_GATHER, st, [&] {
                    hipLaunchKernelGGL(h->gather_kernel_mixed, grid, block, 0, st, vf, nrow, h->face_flux_m.ptr, R,
                                       accumulate);
                });
            } else if (h->strategy == ABL_FV_STRATEGY_TILE) {
                h->timed(ABL_FV_PHASE_FLUX, st, [&] {
                    hipLaunchKernelGGL(h->tile_kernel_mixed, dim3((unsigned)h->tile_n), dim3(h->tile_T), h->tile_lds, st,
                                       vf, h->ph, U, h->recf.ptr, h->gS_f.ptr, h->gY_f.ptr, h->tile_start.ptr,
                                       h->tile_faces.ptr, h->tile_slot.ptr, h->tile_nfm, R, accumulate);
                });
            } else {
                h->timed(ABL_FV_PHASE_FLUX, st, [&] {
                    hipLaunchKernelGGL(h->flux_kernel_mixed, grid, block, 0, st, vf, h->ph, U, h->recf.ptr, R,
                                       accumulate);
                });
            }
        } else if (h->strategy == ABL_FV_STRATEGY_MARCH) {
            h->timed(ABL_FV_PHASE_FLUX, st, [&] {
                hipLaunchKernelGGL(h->march_kernel, dim3((unsigned)h->mr_nseq), dim3(abl::kMarchBlock), h->mr_lds, st,
                                   h->view(), h->ph, U, h->march_view(), R, accumulate);
            });
        } else if (h->strategy == ABL_FV_STRATEGY_FUSED) {
            h->timed(ABL_FV_PHASE_FLUX, st, [&] {
                hipLaunchKernelGGL(h->fused_kernel, dim3((unsigned)h->fu_n), dim3(256), h->fu_lds, st, h->view(), h->ph, U,
                                   h->fused_view(), R, accumulate);
            });
        } else {
        const auto v = h->view();
        if (h->recon || h->visc) {
            h->timed(ABL_FV_PHASE_GRADIENTS, st, [&] {
                hipLaunchKernelGGL(h->grad_kernel, grid, block, 0, st, v, h->ph, U, h->gradC.ptr, h->gradA.ptr,
                                   (void*)h->rec.ptr);
            });
            if (L.ns)
                h->timed(ABL_FV_PHASE_SPECIES_GRADIENTS, st, [&] {
                    hipLaunchKernelGGL(h->species_kernel, grid, block, 0, st, v, h->ph, U, h->gradC.ptr, h->gradA.ptr);
                });
        }
        if (h->strategy == ABL_FV_STRATEGY_TILE) {
            h->timed(ABL_FV_PHASE_FLUX, st, [&] {
                hipLaunchKernelGGL(h->tile_kernel, dim3((unsigned)h->tile_n), dim3(h->tile_T), h->tile_lds, st, v, h->ph,
                                   U, h->gradC.ptr, h->gradA.ptr, h->rec.ptr, h->tile_start.ptr,
                                   h->tile_faces.ptr,
                                   h->tile_slot.ptr, h->tile_nfm, R, accumulate);
            });
        } else if (h->strategy == ABL_FV_STRATEGY_FACE_FLUX) {
            h->timed(ABL_FV_PHASE_FLUX, st, [&] {
                hipLaunchKernelGGL(h->face_flux_kernel, dim3(grid_for(L.F)), block, 0, st, v, h->ph, U, h->gradC.ptr,
                                   h->gradA.ptr, h->rec.ptr, h->face_flux_e.ptr, h->face_flux_s.ptr);
            });
            h->timed(ABL_FV_PHASE_GATHER, st, [&] {
                hipLaunchKernelGGL(h->gather_kernel, grid, block, 0, st, v, L.ns, h->face_flux_e.ptr, h->face_flux_s.ptr,
                                   R, accumulate);
            });
        } else {
            h->timed(ABL_FV_PHASE_FLUX, st, [&] {
                hipLaunchKernelGGL(h->flux_kernel, grid, block, 0, st, v, h->ph, U, h->gradC.ptr, h->gradA.ptr,
                                   h->rec.ptr, R, accumulate);
            });
        }
        }
        if (!accumulate && L.N > L.NI)
            hipLaunchKernelGGL(abl::k_zero_rows, dim3(grid_for((L.N - L.NI) * L.nc)), block, 0, st, R, L.N, L.NI, L.nc);
        HIP_CHECK(hipGetLastError());
    } catch (const std::exception& e) {
        return fail(h, std::string("abl_fv_rhs: ") + e.what());
    }
    return 0;
}

extern "C" int abl_fv_set_strategy(abl_fv_handle h, int32_t strategy) {
    if (!h) return fail(h, "abl_fv_set_strategy: null handle");
    if (strategy != ABL_FV_STRATEGY_AUTO && strategy != ABL_FV_STRATEGY_RECOMPUTE &&
        strategy != ABL_FV_STRATEGY_FACE_FLUX && strategy != ABL_FV_STRATEGY_TILE && strategy != ABL_FV_STRATEGY_FUSED &&
        strategy != ABL_FV_STRATEGY_MARCH)
        return fail(h, "abl_fv_set_strategy: unknown strategy");
    if (strategy == ABL_FV_STRATEGY_FUSED && h->fu_n == 0)
        return fail(h, "abl_fv_set_strategy: the fused strategy is unavailable for this handle (it needs the FP64 form "
                       "without species, the Constant transport model, no limiter or Barth-Jespersen, 2-D cells or "
                       "tetrahedra / hexahedra, prisms, "
                       "pyramids, well-conditioned least-squares stencils and a tile that fits the LDS)" +
                           (h->fused_unavailable.empty() ? std::string() : ": " + h->fused_unavailable));
    if (strategy == ABL_FV_STRATEGY_MARCH && h->mr_nseq == 0)
        return fail(h, "abl_fv_set_strategy: the march strategy is unavailable for this handle (it needs what the fused "
                       "strategy needs and a marching plan whose LDS fits a workgroup)" +
                           (h->march_unavailable.empty() ? std::string() : ": " + h->march_unavailable));
    if (strategy == ABL_FV_STRATEGY_AUTO) strategy = h->auto_strategy();
    if (strategy == ABL_FV_STRATEGY_TILE && h->tile_n == 0)
        return fail(h, "abl_fv_set_strategy: the tile strategy is unavailable for this handle (no tile size fits the "
                       "LDS budget)");
    if (strategy == ABL_FV_STRATEGY_RECOMPUTE && h->form != ABL_FV_FORM_FP64 && h->host.ns > 0)
        return fail(h, "abl_fv_set_strategy: the mixed form transports species with the tile and face-flux "
                       "strategies only");
    try {
        if (strategy == ABL_FV_STRATEGY_FACE_FLUX) h->ensure_face_flux();
    } catch (const std::exception& e) {
        return fail(h, std::string("abl_fv_set_strategy: ") + e.what());
    }
    h->strategy = strategy;
    return 0;
}

extern "C" int abl_fv_set_tile_size(abl_fv_handle h, int32_t cells) {
    if (!h) return fail(h, "abl_fv_set_tile_size: null handle");
    if (cells != 0 && cells != 64 && cells != 128 && cells != 256 && cells != 512)
        return fail(h, "abl_fv_set_tile_size: cells must be 0 (default), 64, 128, 256 or 512");
    try {
        // the host copies of the cell -> face table and the face cells were released at creation
        abl::Layout lay = h->host;
        lay.ell.resize(h->ell.n);
        lay.fcell.resize(h->fcell.n);
        HIP_CHECK(hipMemcpy(lay.ell.data(), h->ell.ptr, h->ell.bytes(), hipMemcpyDeviceToHost));
        HIP_CHECK(hipMemcpy(lay.fcell.data(), h->fcell.ptr, h->fcell.bytes(), hipMemcpyDeviceToHost));
        if (!h->build_tiles(lay, cells))
            return fail(h, "abl_fv_set_tile_size: the tile's face fluxes exceed the workgroup LDS");
    } catch (const std::exception& e) {
        return fail(h, std::string("abl_fv_set_tile_size: ") + e.what());
    }
    return 0;
}

extern "C" int abl_fv_get_strategy(abl_fv_handle h, int32_t* strategy) {
    if (!h || !strategy) return fail(h, "abl_fv_get_strategy: null argument");
    *strategy = h->strategy;
    return 0;
}

extern "C" int abl_fv_set_profiling(abl_fv_handle h, int32_t enable) {
    if (!h) return fail(h, "abl_fv_set_profiling: null handle");
    h->profiling = enable != 0;
    return 0;
}

extern "C" int abl_fv_get_phase_times(abl_fv_handle h, double* ms, int64_t* launches, int32_t reset) {
    if (!h) return fail(h, "abl_fv_get_phase_times: null handle");
    try {
        h->drain_events();
    } catch (const std::exception& e) {
        return fail(h, std::string("abl_fv_get_phase_times: ") + e.what());
    }
    for (int p = 0; p < ABL_FV_N_PHASES; p++) {
        if (ms) ms[p] = h->phase_ms[p];
        if (launches) launches[p] = h->phase_launches[p];
        if (reset) {
            h->phase_ms[p] = 0.0;
            h->phase_launches[p] = 0;
        }
    }
    return 0;
}

namespace {
// theta_bar: per-slot sums of the per-boundary-face contributions (contrib [3 nb]) reduced on the device in a fixed
// order (deterministic, k_theta_reduce) into sum_dev [3 kMaxBcSlots]; only those sums come back, and are added into
// the caller's array for the slots that have boundary faces.
void reduce_theta_into(abl_fv_handle h, hipStream_t st, const double* contrib, double* sum_dev, double* theta_bar) {
    hipLaunchKernelGGL(abl::k_theta_reduce, dim3(3 * abl::kMaxBcSlots), dim3(abl::kThetaBlock), 0, st, h->n_bfaces,
                       h->bslot.ptr, h->slot_kind.ptr, contrib, sum_dev);
    HIP_CHECK(hipGetLastError());
    double sums[3 * abl::kMaxBcSlots];
    HIP_CHECK(hipMemcpyAsync(sums, sum_dev, sizeof(sums), hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
    for (int v : h->bc_slots)
        for (int j = 0; j < 3; j++) theta_bar[3 * v + j] += sums[3 * v + j];
}
void reduce_theta(abl_fv_handle h, hipStream_t st, double* theta_bar) {
    if (!h->theta_sum.ptr) h->theta_sum.alloc((size_t)3 * abl::kMaxBcSlots);
    reduce_theta_into(h, st, h->theta_contrib.ptr, h->theta_sum.ptr, theta_bar);
}

int vjp_second_order(abl_fv_handle h, double* U, const double* Rbar, double* Ubar, double* theta_bar, hipStream_t st) {
    try {
        const abl::Layout& L = h->host;
        const int KM = abl::slot_width(L.K);
        const int64_t NG = L.N - L.NI;
        const int nwp = L.dim == 3 ? abl::AdjFaceRec<3>::pairs(L.ns) : abl::AdjFaceRec<2>::pairs(L.ns);
        if (!h->adjUself.ptr) {
            h->adjUghost.alloc((size_t)L.nc * std::max<int64_t>(NG, 1));
            h->adjUself.alloc((size_t)L.nc * L.NI);
            h->adjContrib.alloc((size_t)KM * L.nc * L.NI);
            h->adjW.alloc((size_t)2 * nwp * ((L.F + 63) / 64 * 64));
            h->adjGL.alloc((size_t)std::max(L.ns, 1) * std::max<int64_t>(NG, 1));
        }
        const dim3 grid(grid_for(L.NI)), block(abl::kBlock);
        if (h->n_bfaces)
            h->timed(ABL_FV_PHASE_GHOST, st, [&] {
                hipLaunchKernelGGL(abl::k_ghost_fill, dim3(grid_for(h->n_bfaces)), block, 0, st, h->n_bfaces,
                                   h->bface.ptr, h->bslot.ptr, h->fcell.ptr, L.F, L.N, L.nc, h->slot_kind.ptr,
                                   h->slot_state.ptr, U);
            });
        const auto v = h->view();
        h->timed(ABL_FV_PHASE_GRADIENTS, st, [&] {
            hipLaunchKernelGGL(h->grad_kernel, grid, block, 0, st, v, h->ph, U, h->gradC.ptr, h->gradA.ptr,
                               (void*)h->rec.ptr);
        });
        if (L.ns)
            h->timed(ABL_FV_PHASE_SPECIES_GRADIENTS, st, [&] {
                hipLaunchKernelGGL(h->species_kernel, grid, block, 0, st, v, h->ph, U, h->gradC.ptr, h->gradA.ptr);
            });
        abl::AdjBufs b{h->adjUghost.ptr, h->adjUself.ptr, h->adjContrib.ptr, h->mirror.ptr, KM,
                       h->adjW.ptr, nwp, h->adjGL.ptr};
        h->timed(ABL_FV_PHASE_VJP, st, [&] {
            hipLaunchKernelGGL(h->adj_faces, dim3(grid_for(L.F)), block, 0, st, v, h->ph, U, Rbar, h->rec.ptr,
                               h->gradC.ptr, h->gradA.ptr, b);
            hipLaunchKernelGGL(h->adj_cells_e, grid, block, 0, st, v, h->ph, U, b);
            if (L.ns) hipLaunchKernelGGL(h->adj_cells_s, grid, block, 0, st, v, h->ph, U, b);
            hipLaunchKernelGGL(h->adj_gather, grid, block, 0, st, v, h->ph, h->ghost_slot.ptr, h->slot_kind.ptr, b,
                               Ubar);
        });
        if (NG > 0) {
            hipLaunchKernelGGL(abl::k_zero_rows, dim3(grid_for(NG * L.nc)), block, 0, st, Ubar, L.N, L.NI, L.nc);
            hipLaunchKernelGGL(abl::k_ghost_adjoint_out, dim3(grid_for(NG)), block, 0, st, NG, L.N, L.NI, L.nc,
                               h->ghost_slot.ptr, h->slot_kind.ptr, h->adjUghost.ptr, Ubar);
        }
        HIP_CHECK(hipGetLastError());
        if (theta_bar && h->n_bfaces && has_inflow(h)) {
            if (!h->theta_contrib.ptr) h->theta_contrib.alloc((size_t)3 * h->n_bfaces);
            hipLaunchKernelGGL(abl::k_adj_theta, dim3(grid_for(h->n_bfaces)), block, 0, st, h->n_bfaces, L.dim, L.ns,
                               L.N, L.NI, h->bface.ptr, h->bslot.ptr, h->fcell.ptr, L.F, h->slot_kind.ptr,
                               h->slot_param.ptr, h->slot_Y.ptr, h->ph, h->adjUghost.ptr, h->theta_contrib.ptr);
            HIP_CHECK(hipGetLastError());
            reduce_theta(h, st, theta_bar);
        }
    } catch (const std::exception& e) {
        return fail(h, std::string("abl_fv_rhs_vjp: ") + e.what());
    }
    return 0;
}
}  // namespace

namespace {
int vjp_first_order(abl_fv_handle h, double* U, const double* Rbar, double* Ubar, double* theta_bar, double* contrib,
                    double* sums, double* rows, hipStream_t st, const char* who);
}  // namespace

extern "C" int abl_fv_rhs_vjp(abl_fv_handle h, double time, double* U, const double* Rbar, double* Ubar,
                              double* theta_bar, abl_fv_stream stream) {
    (void)time;
    if (!h || !U || !Rbar || !Ubar) return fail(h, "abl_fv_rhs_vjp: null argument");
    {
        const std::string bad = unset_boundary(h);
        if (!bad.empty()) return fail(h, "abl_fv_rhs_vjp: " + bad);
    }
    const abl::Layout& L = h->host;
    const int fl = h->phys.flux_calculator;
    if (h->form != ABL_FV_FORM_FP64 ||
        (fl != ABL_FV_FLUX_AUSM && fl != ABL_FV_FLUX_RUSANOV && fl != ABL_FV_FLUX_OFF && fl != ABL_FV_FLUX_NONE))
        return fail(h, "abl_fv_rhs_vjp: supported for the FP64 form and the Ausm / Rusanov / Off flux calculators");
    if (h->recon && h->phys.limiter != ABL_FV_LIMITER_NONE && h->phys.limiter != ABL_FV_LIMITER_BARTH_JESPERSEN)
        return fail(h, "abl_fv_rhs_vjp: with reconstruction the adjoint supports the Barth-Jespersen limiter or none");
    if (h->ph.suth && h->visc)
        return fail(h, "abl_fv_rhs_vjp: the adjoint supports the Constant transport model (Sutherland is forward only)");
    if (h->recon || h->visc) return vjp_second_order(h, U, Rbar, Ubar, theta_bar, (hipStream_t)stream);
    try {
        if (theta_bar && h->n_bfaces && !h->theta_contrib.ptr) h->theta_contrib.alloc((size_t)3 * h->n_bfaces);
        if (theta_bar && h->n_bfaces && !h->theta_sum.ptr) h->theta_sum.alloc((size_t)3 * abl::kMaxBcSlots);
        if (fl != ABL_FV_FLUX_OFF && fl != ABL_FV_FLUX_NONE && L.F > 0 && !h->vjp_rows.ptr)
            h->vjp_rows.alloc((size_t)(2 * (L.dim + 2) + 2 + L.ns) * L.F);
    } catch (const std::exception& e) {
        return fail(h, std::string("abl_fv_rhs_vjp: ") + e.what());
    }
    return vjp_first_order(h, U, Rbar, Ubar, theta_bar, h->theta_contrib.ptr, h->theta_sum.ptr, h->vjp_rows.ptr,
                           (hipStream_t)stream, "abl_fv_rhs_vjp");
}

namespace {
// First-order VJP (ghost fill, ghost rows of Ubar, the VJP kernels, theta_bar) with the theta_bar workspace
// supplied by the caller: contrib [3 n_bfaces] and sums [3 kMaxBcSlots] device doubles (written in full before they
// are read).  rows [2 NE + 2 + ns][F]: the face-centric kernels (each face reversed once); null: the cell-centric kernel
// (each face reversed by both of its cells, no workspace).  Both give the same Ubar bit for bit.
int vjp_first_order(abl_fv_handle h, double* U, const double* Rbar, double* Ubar, double* theta_bar, double* contrib,
                    double* sums, double* rows, hipStream_t st, const char* who) {
    const abl::Layout& L = h->host;
    const int fl = h->phys.flux_calculator;
    try {
        if (h->n_bfaces) {
            h->timed(ABL_FV_PHASE_GHOST, st, [&] {
                hipLaunchKernelGGL(abl::k_ghost_fill, dim3(grid_for(h->n_bfaces)), dim3(abl::kBlock), 0, st, h->n_bfaces,
                                   h->bface.ptr, h->bslot.ptr, h->fcell.ptr, L.F, L.N, L.nc, h->slot_kind.ptr,
                                   h->slot_state.ptr, U);
            });
        }
        const dim3 grid(grid_for(L.NI)), block(abl::kBlock);
        const bool noflux = fl == ABL_FV_FLUX_OFF || fl == ABL_FV_FLUX_NONE;
        if (noflux) {  // a fill kernel, not a memset node, so that a captured call replays correctly
            hipLaunchKernelGGL(abl::k_zero_rows, dim3(grid_for(L.N * L.nc)), block, 0, st, Ubar, L.N, (int64_t)0, L.nc);
        } else {
            const auto v = h->view();
            // ghost rows first: the VJP kernel then writes those of caller-filled ghosts
            if (L.N > L.NI)
                hipLaunchKernelGGL(abl::k_zero_rows, dim3(grid_for((L.N - L.NI) * L.nc)), block, 0, st, Ubar, L.N, L.NI,
                                   L.nc);
            h->timed(ABL_FV_PHASE_VJP, st, [&] {
                if (rows) {
                    hipLaunchKernelGGL(h->vjp_faces, dim3(grid_for(L.F)), block, 0, st, v, h->ph, U, Rbar, rows);
                    hipLaunchKernelGGL(h->vjp_gather, grid, block, 0, st, v, h->ph, Rbar, rows, h->ghost_slot.ptr,
                                       h->slot_kind.ptr, Ubar);
                } else {
                    hipLaunchKernelGGL(h->vjp_kernel, grid, block, 0, st, v, h->ph, U, Rbar, h->ghost_slot.ptr,
                                       h->slot_kind.ptr, Ubar);
                }
            });
        }
        HIP_CHECK(hipGetLastError());
        if (theta_bar && h->n_bfaces && !noflux && has_inflow(h)) {
            const auto v = h->view();
            hipLaunchKernelGGL(h->vjp_theta_kernel, dim3(grid_for(h->n_bfaces)), block, 0, st, v, h->ph, U, Rbar,
                               h->n_bfaces, h->bface.ptr, h->bslot.ptr, h->slot_kind.ptr, h->slot_param.ptr,
                               h->slot_Y.ptr, contrib);
            HIP_CHECK(hipGetLastError());
            reduce_theta_into(h, st, contrib, sums, theta_bar);
        }
    } catch (const std::exception& e) {
        return fail(h, std::string(who) + ": " + e.what());
    }
    return 0;
}
}  // namespace

extern "C" int abl_fv_face_interpolate(abl_fv_handle h, const double* U, double* face_sol, double* face_sol_grad,
                                       double* face_aux, double* face_aux_grad, abl_fv_stream stream) {
    if (!h || !U) return fail(h, "abl_fv_face_interpolate: null argument");
    if (h->form != ABL_FV_FORM_FP64) return fail(h, "abl_fv_face_interpolate: FP64 form only");
    if (!h->visc)
        return fail(h, "abl_fv_face_interpolate: the handle has no face interpolant (no diffusion terms: k, mu and "
                       "diff are all zero)");
    try {
        const abl::Layout& L = h->host;
        hipStream_t st = (hipStream_t)stream;
        const auto v = h->view();
        const auto kern = L.dim == 3 ? abl::pick_face_interpolate<3>(L.K) : abl::pick_face_interpolate<2>(L.K);
        hipLaunchKernelGGL(kern, dim3(grid_for(L.F)), dim3(abl::kBlock), 0, st, v, h->ph, U, L.nc, face_sol,
                           face_sol_grad, face_aux, face_aux_grad);
        HIP_CHECK(hipGetLastError());
    } catch (const std::exception& e) {
        return fail(h, std::string("abl_fv_face_interpolate: ") + e.what());
    }
    return 0;
}

extern "C" int abl_fv_pack_aos(abl_fv_handle h, const double* aos, double* soa, abl_fv_stream stream) {
    if (!h || !aos || !soa) return fail(h, "abl_fv_pack_aos: null argument");
    const abl::Layout& L = h->host;
    hipLaunchKernelGGL(abl::k_pack, dim3(grid_for(L.N * L.nc)), dim3(abl::kBlock), 0, (hipStream_t)stream, aos, soa, L.N, L.nc, 1);
    if (hipGetLastError() != hipSuccess) return fail(h, "abl_fv_pack_aos: launch failed");
    return 0;
}

extern "C" int abl_fv_unpack_aos(abl_fv_handle h, const double* soa, double* aos, abl_fv_stream stream) {
    if (!h || !aos || !soa) return fail(h, "abl_fv_unpack_aos: null argument");
    const abl::Layout& L = h->host;
    hipLaunchKernelGGL(abl::k_pack, dim3(grid_for(L.N * L.nc)), dim3(abl::kBlock), 0, (hipStream_t)stream, soa, aos, L.N, L.nc, 0);
    if (hipGetLastError() != hipSuccess) return fail(h, "abl_fv_unpack_aos: launch failed");
    return 0;
}

namespace {
// CFL step size of state U (device), returned on the host (synchronises the stream).
double cfl_dt_impl(abl_fv_handle h, const double* U, hipStream_t st) {
    const abl::Layout& L = h->host;
    const unsigned nb = grid_for(L.NI);
    // NavierStokesTransport's CFL uses pgsAlpha = 1; the compact process divides a by the PGS alpha
    // (compactCompressibleNSSpeciesTransport.cpp:305-339)
    const double alpha = h->ph.compact ? h->ph.pgs_alpha : 1.0;
    hipLaunchKernelGGL(abl::k_cfl_partial, dim3(nb), dim3(abl::kBlock), 0, st, U, h->rmin.ptr, L.N, L.NI, L.dim, h->ph,
                       h->phys.cfl, alpha, h->cfl_partial.ptr);
    hipLaunchKernelGGL(abl::k_min_final, dim3(1), dim3(abl::kBlock), 0, st, h->cfl_partial.ptr, (int64_t)nb,
                       h->cfl_out.ptr);
    HIP_CHECK(hipGetLastError());
    double dt = 0.0;
    HIP_CHECK(hipMemcpyAsync(&dt, h->cfl_out.ptr, sizeof(double), hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
    return dt;
}
// ComputePhysicsTimeSteps: {cfl, conduction, viscous} on the host and the mask of registered functions
// (synchronises the stream).
int time_steps_impl(abl_fv_handle h, const double* U, double* dt, hipStream_t st) {
    const abl::Layout& L = h->host;
    const double fc = h->phys.conduction_stability_factor, fvs = h->phys.viscous_stability_factor;
    const bool has_transport = h->visc;  // a transport model with an enabled property
    const int reg = (h->phys.flux_calculator != ABL_FV_FLUX_NONE ? 1 : 0) | (has_transport && fc > 0.0 ? 2 : 0) |
                    (has_transport && fvs > 0.0 ? 4 : 0);
    dt[0] = dt[1] = dt[2] = 1e10;
    if (reg & 1) dt[0] = cfl_dt_impl(h, U, st);
    if (reg & 6) {
        const unsigned nb = grid_for(L.NI);
        hipLaunchKernelGGL(abl::k_diffusion_dt_partial, dim3(nb), dim3(abl::kBlock), 0, st, U, h->rmin.ptr, L.N, L.NI,
                           L.dim, h->ph, fc, fvs, h->dt_partial.ptr);
        for (int j = 0; j < 2; j++)
            hipLaunchKernelGGL(abl::k_min_final, dim3(1), dim3(abl::kBlock), 0, st, h->dt_partial.ptr + (size_t)j * nb,
                               (int64_t)nb, h->cfl_out.ptr + 1 + j);
        HIP_CHECK(hipGetLastError());
        double v[2] = {0.0, 0.0};
        HIP_CHECK(hipMemcpyAsync(v, h->cfl_out.ptr + 1, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipStreamSynchronize(st));
        if (reg & 2) dt[1] = v[0];
        if (reg & 4) dt[2] = v[1];
    }
    return reg;
}
// ComputePhysicsTimeStep: the minimum over the registered functions, Constants::large when none.
double physics_dt_impl(abl_fv_handle h, const double* U, hipStream_t st) {
    double dt[3];
    const int reg = time_steps_impl(h, U, dt, st);
    double m = 1e10;
    for (int j = 0; j < 3; j++)
        if (reg & (1 << j)) m = std::min(m, dt[j]);
    return m;
}
void axpy_impl(abl_fv_handle h, double dt, const double* R, double* U, hipStream_t st) {
    const abl::Layout& L = h->host;
    hipLaunchKernelGGL(abl::k_axpy, dim3(grid_for(L.NI), (unsigned)L.nc), dim3(abl::kBlock), 0, st, dt, R, U, L.N, L.NI);
    HIP_CHECK(hipGetLastError());
}
}  // namespace

extern "C" int abl_fv_cfl_dt(abl_fv_handle h, const double* U, double* dt, abl_fv_stream stream) {
    if (!h || !U || !dt) return fail(h, "abl_fv_cfl_dt: null argument");
    if (!(h->phys.cfl > 0.0)) return fail(h, "abl_fv_cfl_dt: the physics descriptor has no positive cfl");
    try {
        *dt = cfl_dt_impl(h, U, (hipStream_t)stream);
    } catch (const std::exception& e) {
        return fail(h, std::string("abl_fv_cfl_dt: ") + e.what());
    }
    return 0;
}

extern "C" int abl_fv_time_steps(abl_fv_handle h, const double* U, double* dt, int32_t* registered,
                                 abl_fv_stream stream) {
    if (!h || !U || !dt || !registered) return fail(h, "abl_fv_time_steps: null argument");
    try {
        *registered = time_steps_impl(h, U, dt, (hipStream_t)stream);
    } catch (const std::exception& e) {
        return fail(h, std::string("abl_fv_time_steps: ") + e.what());
    }
    return 0;
}

extern "C" int abl_fv_physics_dt(abl_fv_handle h, const double* U, double* dt, abl_fv_stream stream) {
    if (!h || !U || !dt) return fail(h, "abl_fv_physics_dt: null argument");
    try {
        *dt = physics_dt_impl(h, U, (hipStream_t)stream);
    } catch (const std::exception& e) {
        return fail(h, std::string("abl_fv_physics_dt: ") + e.what());
    }
    return 0;
}

extern "C" int abl_fv_axpy(abl_fv_handle h, double dt, const double* R, double* U, abl_fv_stream stream) {
    if (!h || !R || !U) return fail(h, "abl_fv_axpy: null argument");
    try {
        axpy_impl(h, dt, R, U, (hipStream_t)stream);
    } catch (const std::exception& e) {
        return fail(h, std::string("abl_fv_axpy: ") + e.what());
    }
    return 0;
}

extern "C" int abl_fv_set_pgs(abl_fv_handle h, const abl_fv_pgs_desc* pgs) {
    if (!h) return fail(h, "abl_fv_set_pgs: null handle");
    if (!pgs) {
        h->pgs_on = false;
        h->ph.pgs_alpha = 1.0;
        return 0;
    }
    try {
        h->pgs = *pgs;
        // constructor defaults, pressureGradientScaling.cpp:10-11
        h->pgs.max_alpha_allowed = pgs->max_alpha_allowed >= 1.0 ? pgs->max_alpha_allowed : 100.0;
        h->pgs.max_delta_pressure_fac = pgs->max_delta_pressure_fac > 0.0 ? pgs->max_delta_pressure_fac : 0.05;
        if (!h->pgs_out.ptr) {
            h->pgs_partial.alloc((size_t)std::max<int64_t>(grid_for(h->host.NI), 1) * abl::kPgsN);
            h->pgs_out.alloc(abl::kPgsN);
        }
        h->ph.pgs_alpha = pgs->alpha_init;
        h->pgs_max_mach = 0.0;
        h->pgs_on = true;
    } catch (const std::exception& e) {
        return fail(h, std::string("abl_fv_set_pgs: ") + e.what());
    }
    return 0;
}

extern "C" int abl_fv_pgs_update(abl_fv_handle h, const double* U, double dt, double* alpha, double* max_mach,
                                 abl_fv_stream stream) {
    if (!h || !U) return fail(h, "abl_fv_pgs_update: null argument");
    if (!h->pgs_on) return fail(h, "abl_fv_pgs_update: no PressureGradientScaling attached (abl_fv_set_pgs)");
    try {
        const abl::Layout& L = h->host;
        hipStream_t st = (hipStream_t)stream;
        const unsigned nb = grid_for(L.NI);
        hipLaunchKernelGGL(abl::k_pgs_partial, dim3(nb), dim3(abl::kBlock), 0, st, U, L.N, L.NI, L.dim, h->ph,
                           h->pgs_partial.ptr);
        hipLaunchKernelGGL(abl::k_pgs_final, dim3(1), dim3(abl::kBlock), 0, st, h->pgs_partial.ptr, (int64_t)nb,
                           h->pgs_out.ptr);
        HIP_CHECK(hipGetLastError());
        double r[abl::kPgsN];
        HIP_CHECK(hipMemcpyAsync(r, h->pgs_out.ptr, sizeof(r), hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipStreamSynchronize(st));
        // the alpha update of pressureGradientScaling.cpp:106-131 (PetscMin / PetscMax comparison order kept)
        auto pmin = [](double a, double b) { return a < b ? a : b; };
        auto pmax = [](double a, double b) { return a < b ? b : a; };
        const double kMaxMachAllowed = 0.7, kMaxAlphaChange = 0.10;  // pressureGradientScaling.hpp:32,37
        const double pRef = r[0] / (double)L.NI;
        const double maxDeltaP = pmax(std::fabs(r[2] - pRef), std::fabs(r[1] - pRef));
        const double cMax = r[3];
        const double machMax = r[4];
        // min over cells of maxMachAllowed / Mach == maxMachAllowed / max Mach (division is monotone)
        const double alphaMax = machMax > 0.0 ? pmin(DBL_MAX, kMaxMachAllowed / machMax) : DBL_MAX;
        const double alphaOld = h->ph.pgs_alpha;
        const double term1 = 0.5e+0 * dt * cMax / h->pgs.domain_length;
        double a = alphaOld + term1 * (std::sqrt(h->pgs.max_delta_pressure_fac * pRef / (maxDeltaP + 1E-30)) - 1.0);
        a = pmin(a, (1. + kMaxAlphaChange) * alphaOld);
        a = pmin(a, alphaMax);
        a = pmin(a, h->pgs.max_alpha_allowed);
        a = pmax(a, 1.e+0);
        h->ph.pgs_alpha = a;
        h->pgs_max_mach = machMax;
        if (alpha) *alpha = a;
        if (max_mach) *max_mach = machMax;
    } catch (const std::exception& e) {
        return fail(h, std::string("abl_fv_pgs_update: ") + e.what());
    }
    return 0;
}

extern "C" int abl_fv_get_pgs_alpha(abl_fv_handle h, double* alpha) {
    if (!h || !alpha) return fail(h, "abl_fv_get_pgs_alpha: null argument");
    *alpha = h->ph.pgs_alpha;
    return 0;
}

extern "C" int abl_fv_set_pgs_alpha(abl_fv_handle h, double alpha) {
    if (!h) return fail(h, "abl_fv_set_pgs_alpha: null handle");
    if (!h->pgs_on) return fail(h, "abl_fv_set_pgs_alpha: no PressureGradientScaling attached (abl_fv_set_pgs)");
    h->ph.pgs_alpha = alpha;
    return 0;
}

extern "C" int abl_fv_ts_euler(abl_fv_handle h, double* U, double* R, double t0, double max_time, double dt0,
                               double safety, int64_t max_steps, double* t_out, int64_t* steps_out,
                               abl_fv_stream stream) {
    if (!h || !U || !R) return fail(h, "abl_fv_ts_euler: null argument");
    if (!(h->phys.cfl > 0.0) || !(safety > 0.0) || !(dt0 > 0.0))
        return fail(h, "abl_fv_ts_euler: cfl, safety and dt0 must be positive");
    {
        const std::string bad = unset_boundary(h);
        if (!bad.empty()) return fail(h, "abl_fv_ts_euler: " + bad);
    }
    hipStream_t st = (hipStream_t)stream;
    double t = t0;
    int64_t steps = 0;
    try {
        double dt = std::min(dt0, safety * physics_dt_impl(h, U, st));  // AdaptPhysics::AdaptInitializer
        while (t < max_time && steps < max_steps) {
            if (!(dt > 0.0)) throw std::runtime_error("non-positive or NaN step size at t = " + std::to_string(t));
            // TS pre-step: PressureGradientScaling::UpdatePreconditioner with the step about to be taken
            // (registered by PressureGradientScaling::Setup, pressureGradientScaling.cpp:144-146)
            if (h->pgs_on)
                if (int rc = abl_fv_pgs_update(h, U, dt, nullptr, nullptr, stream)) return rc;
            if (int rc = abl_fv_rhs(h, t, U, R, 0, stream)) return rc;
            // TSStep_Euler asks the adaptor for the next step while the solution still holds U_n
            double next = safety * physics_dt_impl(h, U, st);
            axpy_impl(h, dt, R, U, st);
            const double t_end = t + dt, remaining = max_time - t_end;
            // TSAdaptChoose with ts_exact_final_time matchstep (matchstep factors 0.01, 2)
            if (t_end < max_time && t_end + next > max_time) next = remaining;
            if (t_end < max_time && t_end + next < max_time && next * 2.0 > remaining) next = remaining / 2.0;
            if (t_end < max_time && t_end + next < max_time && next * 1.01 > remaining) next = remaining;
            t = t_end;
            dt = next;
            steps++;
        }
        HIP_CHECK(hipStreamSynchronize(st));
    } catch (const std::exception& e) {
        return fail(h, std::string("abl_fv_ts_euler: ") + e.what());
    }
    if (t_out) *t_out = t;
    if (steps_out) *steps_out = steps;
    return 0;
}

extern "C" int abl_fv_check_state(abl_fv_handle h, const double* U, int64_t* n_bad, abl_fv_stream stream) {
    if (!h || !U || !n_bad) return fail(h, "abl_fv_check_state: null argument");
    try {
        const abl::Layout& L = h->host;
        hipStream_t st = (hipStream_t)stream;
        HIP_CHECK(hipMemsetAsync(h->counter.ptr, 0, sizeof(unsigned long long), st));
        hipLaunchKernelGGL(abl::k_check_state, dim3(grid_for(L.NI)), dim3(abl::kBlock), 0, st, U, L.N, L.NI, L.nc, h->counter.ptr);
        HIP_CHECK(hipGetLastError());
        unsigned long long v = 0;
        HIP_CHECK(hipMemcpyAsync(&v, h->counter.ptr, sizeof(v), hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipStreamSynchronize(st));
        *n_bad = (int64_t)v;
    } catch (const std::exception& e) {
        return fail(h, std::string("abl_fv_check_state: ") + e.what());
    }
    return 0;
}

// ============================================================================================ allocation-free VJP
//
// abl_fv_rhs_vjp_noalloc: the first-order VJP with its theta_bar workspace (per-boundary-face contributions
// [3 nb] and per-slot sums [3 kMaxBcSlots]) in a file-scope device array instead of memory the library allocates,
// for callers that cannot allocate (HIP graph capture, serving loops).  Each GPU holds its own copy of the array
// until the process ends; calls on one GPU share it, so calls that overlap on one GPU race (serialise them).  Every
// word of it that a call reads is written earlier in the same call.
namespace {
constexpr int64_t kNoAllocMaxBoundaryFaces = 1 << 18;
constexpr int64_t kNoAllocScratchDoubles = 3 * kNoAllocMaxBoundaryFaces + 3 * abl::kMaxBcSlots;
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
double* g_scratch_dev[kMaxDevices] = {nullptr};  // address of `scratch` on each device (hipGetSymbolAddress)
}  // namespace

static __device__ __align__(256) double scratch[kNoAllocScratchDoubles];

namespace {
// The current device's address of `scratch`, resolved once per device with that device current.
double* scratch_for_current_device() {
    int dev = 0;
    HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) throw std::runtime_error("device ordinal out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_dev[dev] = static_cast<double*>(p);
    }
    return g_scratch_dev[dev];
}
}  // namespace

namespace abl {
// Called by abl_fv_create on the handle's device: resolves the scratch address at set-up.
void resolve_noalloc_scratch() { (void)scratch_for_current_device(); }
}  // namespace abl

extern "C" int abl_fv_rhs_vjp_noalloc(abl_fv_handle h, double time, double* U, const double* Rbar, double* Ubar,
                                      double* theta_bar, abl_fv_stream stream) {
    (void)time;
    if (!h || !U || !Rbar || !Ubar) return fail(h, "abl_fv_rhs_vjp_noalloc: null argument");
    {
        const std::string bad = unset_boundary(h);
        if (!bad.empty()) return fail(h, "abl_fv_rhs_vjp_noalloc: " + bad);
    }
    const int fl = h->phys.flux_calculator;
    if (h->form != ABL_FV_FORM_FP64 ||
        (fl != ABL_FV_FLUX_AUSM && fl != ABL_FV_FLUX_RUSANOV && fl != ABL_FV_FLUX_OFF && fl != ABL_FV_FLUX_NONE))
        return fail(h, "abl_fv_rhs_vjp_noalloc: supported for the FP64 form and the Ausm / Rusanov / Off flux "
                       "calculators");
    if (h->recon || h->visc)
        return fail(h, "abl_fv_rhs_vjp_noalloc: first-order inviscid handles only (the second-order adjoint needs "
                       "per-cell workspace: use abl_fv_rhs_vjp)");
    if (h->n_bfaces > kNoAllocMaxBoundaryFaces)
        return fail(h, "abl_fv_rhs_vjp_noalloc: " + std::to_string(h->n_bfaces) + " boundary faces exceed the maximum " +
                           std::to_string(kNoAllocMaxBoundaryFaces));
    double* sc = nullptr;
    try {
        sc = scratch_for_current_device();
    } catch (const std::exception& e) {
        return fail(h, std::string("abl_fv_rhs_vjp_noalloc: ") + e.what());
    }
    return vjp_first_order(h, U, Rbar, Ubar, theta_bar, sc, sc + 3 * kNoAllocMaxBoundaryFaces, nullptr,
                           (hipStream_t)stream, "abl_fv_rhs_vjp_noalloc");
}
