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
#include <stdexcept>
#include <string>
#include <vector>

#include "fv_device.hpp"
#include "fv_internal.hpp"
#include "fv_layout.hpp"

namespace abl {

namespace {
thread_local std::string g_last_error;
}
void set_global_error(const std::string& msg) { g_last_error = msg; }

constexpr int kBlock = 256;

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
    for (int p = 0; p < nc; p++) {
        double v = kind == ABL_FV_BC_INFLOW ? slot_state[s * nc + p] : U[(int64_t)p * N + interior];
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

// Gradients of the euler field (limited) and of the aux T/velocity (unlimited), one interior cell per lane.
template <int DIM, bool RECON, bool VISC, int LIM>
__global__ __launch_bounds__(kBlock) void k_gradients(MeshView<double> m, Phys ph, const double* __restrict__ U,
                                                      double* __restrict__ gradC, double* __restrict__ gradA) {
    constexpr int NE = DIM + 2;
    constexpr int NA = DIM + 1;
    const int64_t c = xcd_cell_index();
    if (c >= m.NI) return;
    const int64_t N = m.N, F = m.F, NI = m.NI;
    const double* __restrict__ wa_arr = m.fwa ? m.fwa : m.fw;
    double uc[NE], ac[NA];
#pragma unroll
    for (int p = 0; p < NE; p++) uc[p] = U[p * N + c];
    if (VISC) aux_of<DIM>(uc, ph, ac);
    double g[NE][DIM], ga[NA][DIM], umin[NE], umax[NE];
#pragma unroll
    for (int p = 0; p < NE; p++) {
        umin[p] = uc[p];
        umax[p] = uc[p];
#pragma unroll
        for (int d = 0; d < DIM; d++) g[p][d] = 0.0;
    }
#pragma unroll
    for (int q = 0; q < NA; q++)
#pragma unroll
        for (int d = 0; d < DIM; d++) ga[q][d] = 0.0;

    for (int k = 0; k < m.K; k++) {
        const int32_t e = m.ell[k * NI + c];
        if (e < 0) break;
        const int64_t f = e >> 1;
        const int s = e & 1;
        const int64_t n = m.fcell[(1 - s) * F + f];
        const bool bnd = n >= NI;
        double un[NE];
#pragma unroll
        for (int p = 0; p < NE; p++) un[p] = U[p * N + n];
        if (RECON) {
            if (LIM == kLimBJ) {
#pragma unroll
                for (int p = 0; p < NE; p++) {
                    umin[p] = fmin(umin[p], un[p]);
                    umax[p] = fmax(umax[p], un[p]);
                }
            }
            if (!bnd) {
                double w[2][DIM];
                load_face_pair<DIM>(m.fw, f, w);
#pragma unroll
                for (int p = 0; p < NE; p++) {
                    // delta = u_R - u_L; the left cell adds w.delta, the right cell subtracts it
                    if (s == 0) {
                        const double delta = un[p] - uc[p];
#pragma unroll
                        for (int d = 0; d < DIM; d++) g[p][d] += w[0][d] * delta;
                    } else {
                        const double delta = uc[p] - un[p];
#pragma unroll
                        for (int d = 0; d < DIM; d++) g[p][d] -= w[1][d] * delta;
                    }
                }
            }
        }
        if (VISC) {
            double an[NA], w[2][DIM];
            aux_of<DIM>(un, ph, an);
            load_face_pair<DIM>(wa_arr, f, w);
#pragma unroll
            for (int q = 0; q < NA; q++) {
                const double delta = an[q] - ac[q];
#pragma unroll
                for (int d = 0; d < DIM; d++) ga[q][d] += w[s][d] * delta;
            }
        }
    }

    if (RECON && LIM != kLimNone) {
        double phi[NE];
        if (LIM == kLimBJ) {
            // phi = min(1, (u_max - u_c)/max_f D_f+, (u_min - u_c)/min_f D_f-), D_f = grad.(x_f - x_c): the per-face
            // minimum collapsed onto the extreme projections (division is monotone).
            double dpos[NE], dneg[NE];
#pragma unroll
            for (int p = 0; p < NE; p++) {
                dpos[p] = 0.0;
                dneg[p] = 0.0;
            }
            for (int k = 0; k < m.K; k++) {
                const int32_t e = m.ell[k * NI + c];
                if (e < 0) break;
                const int64_t f = e >> 1;
                const int s = e & 1;
                double x[2][DIM];
                load_face_pair<DIM>(m.fdx, f, x);
#pragma unroll
                for (int p = 0; p < NE; p++) {
                    double D = 0.0;
#pragma unroll
                    for (int d = 0; d < DIM; d++) D += g[p][d] * x[s][d];
                    dpos[p] = fmax(dpos[p], D);
                    dneg[p] = fmin(dneg[p], D);
                }
            }
#pragma unroll
            for (int p = 0; p < NE; p++) {
                double v = 1.0;
                if (dpos[p] > 0.0) v = fmin(v, (umax[p] - uc[p]) / dpos[p]);
                if (dneg[p] < 0.0) v = fmin(v, (umin[p] - uc[p]) / dneg[p]);
                phi[p] = v;
            }
        } else {
            // DMPlexApplyLimiter_Internal (cellInterpolant.cpp:282-328): f = 1/2 (u_n - u_c) / (grad.(x_n - x_c))
#pragma unroll
            for (int p = 0; p < NE; p++) phi[p] = DBL_MAX;
            for (int k = 0; k < m.K; k++) {
                const int32_t e = m.ell[k * NI + c];
                if (e < 0) break;
                const int64_t f = e >> 1;
                const int s = e & 1;
                const int64_t n = m.fcell[(1 - s) * F + f];
                double x[2][DIM], v[DIM];
                load_face_pair<DIM>(m.fdx, f, x);
#pragma unroll
                for (int d = 0; d < DIM; d++) v[d] = x[s][d] - x[1 - s][d];
#pragma unroll
                for (int p = 0; p < NE; p++) {
                    double denom = 0.0;
#pragma unroll
                    for (int d = 0; d < DIM; d++) denom += g[p][d] * v[d];
                    const double flim = 0.5 * (U[p * N + n] - uc[p]) / denom;
                    phi[p] = pmin(phi[p], limiter_phi<double>(ph.limiter, flim));
                }
            }
        }
        bool cancel = false;
#pragma unroll
        for (int p = 0; p < NE; p++)
            if (phi[p] == 0)
#pragma unroll
                for (int d = 0; d < DIM; d++)
                    if (fabs(g[p][d]) > ph.max_lim_grad) cancel = true;
#pragma unroll
        for (int p = 0; p < NE; p++)
#pragma unroll
            for (int d = 0; d < DIM; d++) g[p][d] *= cancel ? 0.0 : phi[p];
    }
    if (RECON)
#pragma unroll
        for (int p = 0; p < NE; p++)
#pragma unroll
            for (int d = 0; d < DIM; d++) gradC[(p * DIM + d) * N + c] = g[p][d];
    if (VISC)
#pragma unroll
        for (int q = 0; q < NA; q++)
#pragma unroll
            for (int d = 0; d < DIM; d++) gradA[(q * DIM + d) * N + c] = ga[q][d];
}

// densityYi gradients (limited, one field) and Yi aux gradients (unlimited).
template <int DIM>
__global__ __launch_bounds__(kBlock) void k_species_gradients(MeshView<double> m, Phys ph, const double* __restrict__ U,
                                                              double* __restrict__ gradC, double* __restrict__ gradA) {
    constexpr int NE = DIM + 2;
    const int64_t c = xcd_cell_index();
    if (c >= m.NI) return;
    const int64_t N = m.N, F = m.F, NI = m.NI;
    const bool recon = ph.recon != 0;
    const bool aux = ph.sdiff != 0;
    const double* __restrict__ wa_arr = m.fwa ? m.fwa : m.fw;
    const double irho_c = 1.0 / U[kRho * N + c];
    bool cancel = false;
    for (int sp = 0; sp < ph.ns; sp++) {
        const int64_t comp = NE + sp;
        const double uc = U[comp * N + c];
        const double yc = uc * irho_c;
        double g[DIM], ga[DIM], umin = uc, umax = uc;
#pragma unroll
        for (int d = 0; d < DIM; d++) {
            g[d] = 0.0;
            ga[d] = 0.0;
        }
        for (int k = 0; k < m.K; k++) {
            const int32_t e = m.ell[k * NI + c];
            if (e < 0) break;
            const int64_t f = e >> 1;
            const int s = e & 1;
            const int64_t n = m.fcell[(1 - s) * F + f];
            const double un = U[comp * N + n];
            if (recon) {
                umin = fmin(umin, un);
                umax = fmax(umax, un);
                if (n < NI) {
                    double w[2][DIM];
                    load_face_pair<DIM>(m.fw, f, w);
                    if (s == 0) {
                        const double delta = un - uc;
#pragma unroll
                        for (int d = 0; d < DIM; d++) g[d] += w[0][d] * delta;
                    } else {
                        const double delta = uc - un;
#pragma unroll
                        for (int d = 0; d < DIM; d++) g[d] -= w[1][d] * delta;
                    }
                }
            }
            if (aux) {
                double w[2][DIM];
                load_face_pair<DIM>(wa_arr, f, w);
                const double delta = un / U[kRho * N + n] - yc;
#pragma unroll
                for (int d = 0; d < DIM; d++) ga[d] += w[s][d] * delta;
            }
        }
        if (recon) {
            double phi = DBL_MAX;
            if (ph.limiter == ABL_FV_LIMITER_BARTH_JESPERSEN) {
                double dpos = 0.0, dneg = 0.0;
                for (int k = 0; k < m.K; k++) {
                    const int32_t e = m.ell[k * NI + c];
                    if (e < 0) break;
                    const int64_t f = e >> 1;
                    const int s = e & 1;
                    double x[2][DIM];
                    load_face_pair<DIM>(m.fdx, f, x);
                    double D = 0.0;
#pragma unroll
                    for (int d = 0; d < DIM; d++) D += g[d] * x[s][d];
                    dpos = fmax(dpos, D);
                    dneg = fmin(dneg, D);
                }
                phi = 1.0;
                if (dpos > 0.0) phi = fmin(phi, (umax - uc) / dpos);
                if (dneg < 0.0) phi = fmin(phi, (umin - uc) / dneg);
            } else if (ph.limiter != ABL_FV_LIMITER_NONE) {
                for (int k = 0; k < m.K; k++) {
                    const int32_t e = m.ell[k * NI + c];
                    if (e < 0) break;
                    const int64_t f = e >> 1;
                    const int s = e & 1;
                    const int64_t n = m.fcell[(1 - s) * F + f];
                    double x[2][DIM];
                    load_face_pair<DIM>(m.fdx, f, x);
                    double denom = 0.0;
#pragma unroll
                    for (int d = 0; d < DIM; d++) denom += g[d] * (x[s][d] - x[1 - s][d]);
                    const double flim = 0.5 * (U[comp * N + n] - uc) / denom;
                    phi = pmin(phi, limiter_phi<double>(ph.limiter, flim));
                }
            } else {
                phi = 1.0;
            }
            if (phi == 0)
#pragma unroll
                for (int d = 0; d < DIM; d++)
                    if (fabs(g[d]) > ph.max_lim_grad) cancel = true;
#pragma unroll
            for (int d = 0; d < DIM; d++) gradC[(comp * DIM + d) * N + c] = g[d] * phi;
        }
        if (aux)
#pragma unroll
            for (int d = 0; d < DIM; d++) gradA[((DIM + 1 + sp) * DIM + d) * N + c] = ga[d];
    }
    if (cancel)  // the maxLimGrad rule zeroes the whole densityYi field of the cell (cellInterpolant.cpp:459-481)
        for (int sp = 0; sp < ph.ns; sp++)
#pragma unroll
            for (int d = 0; d < DIM; d++) gradC[((NE + sp) * DIM + d) * N + c] = 0.0;
}

// Face fluxes gathered per cell.  One lane per interior cell; both cells of a face evaluate its flux with the
// reference orientation, so the gathered residual is conservative and order-deterministic.
template <int DIM, bool RECON, bool VISC, bool SPECIES>
__global__ __launch_bounds__(kBlock) void k_flux(MeshView<double> m, Phys ph, const double* __restrict__ U,
                                                 const double* __restrict__ gradC, const double* __restrict__ gradA,
                                                 double* __restrict__ R, int accumulate) {
    constexpr int NE = DIM + 2;
    constexpr int NA = DIM + 1;
    constexpr int MAXK = SPECIES ? kMaxFacesPerCell : 1;
    const int64_t c = xcd_cell_index();
    if (c >= m.NI) return;
    const int64_t N = m.N, F = m.F, NI = m.NI;
    const double* __restrict__ wa_arr = m.fwa ? m.fwa : m.fw;

    double uc[NE], gc[NE][DIM], ac[NA], gac[NA][DIM];
#pragma unroll
    for (int p = 0; p < NE; p++) {
        uc[p] = U[p * N + c];
#pragma unroll
        for (int d = 0; d < DIM; d++) gc[p][d] = RECON ? gradC[(p * DIM + d) * N + c] : 0.0;
    }
    if (VISC) {
        aux_of<DIM>(uc, ph, ac);
#pragma unroll
        for (int q = 0; q < NA; q++)
#pragma unroll
            for (int d = 0; d < DIM; d++) gac[q][d] = gradA[(q * DIM + d) * N + c];
    }
    double acc[NE];
#pragma unroll
    for (int p = 0; p < NE; p++) acc[p] = 0.0;
    // per-face coefficients for the species pass: contribution = cL * Y_L + cR * Y_R (+ diffusion)
    double sp_cl[MAXK], sp_cr[MAXK], sp_cd[MAXK];
    int sp_face[MAXK];
    int nk = 0;

    for (int k = 0; k < m.K; k++) {
        const int32_t e = m.ell[k * NI + c];
        if (e < 0) break;
        const int64_t f = e >> 1;
        const int s = e & 1;
        const int64_t n = m.fcell[(1 - s) * F + f];
        const bool ghost = n >= NI;
        double nh[3], area;
        load_face_na(m.fna, f, nh, area);
        double nhat[DIM];
#pragma unroll
        for (int d = 0; d < DIM; d++) nhat[d] = nh[d];
        double un[NE];
#pragma unroll
        for (int p = 0; p < NE; p++) un[p] = U[p * N + n];
        // ProjectToFace for both sides (ghost gradients are zero)
        double uself[NE], unbr[NE];
        if (RECON) {
            double x[2][DIM];
            load_face_pair<DIM>(m.fdx, f, x);
#pragma unroll
            for (int p = 0; p < NE; p++) {
                double a = 0.0, b = 0.0;
#pragma unroll
                for (int d = 0; d < DIM; d++) {
                    a += gc[p][d] * x[s][d];
                    if (!ghost) b += gradC[(p * DIM + d) * N + n] * x[1 - s][d];
                }
                uself[p] = uc[p] + a;
                unbr[p] = un[p] + b;
            }
        } else {
#pragma unroll
            for (int p = 0; p < NE; p++) {
                uself[p] = uc[p];
                unbr[p] = un[p];
            }
        }
        const double* uL = s == 0 ? uself : unbr;
        const double* uR = s == 0 ? unbr : uself;
        double flux[NE];
#pragma unroll
        for (int p = 0; p < NE; p++) flux[p] = 0.0;
        double cl = 0.0, cr = 0.0;
        if (ph.flux == ABL_FV_FLUX_RUSANOV) {
            double unL, unR, lam;
            rusanov_flux<double, DIM>(uL, uR, nhat, area, ph, flux, unL, unR, lam);
            cl = area * (0.5 * unL + 0.5 * lam);
            cr = area * (0.5 * unR - 0.5 * lam);
        } else if (ph.flux != ABL_FV_FLUX_NONE) {
            double mA, rl, rr;
            int right;
            advection_flux<double, DIM>(ph.flux, uL, uR, nhat, area, ph, flux, mA, right, rl, rr);
            if (right)
                cr = mA / rr;
            else
                cl = mA / rl;
        }
        double cd = 0.0;
        if (VISC) {
            // face interpolant: value = 1/2 (a_L + a_R), gradient = 1/2 (G_L + G_R) of unlimited LSQ gradients
            double an[NA], gan[NA][DIM];
            aux_of<DIM>(un, ph, an);
            if (ghost) {
                double w[2][DIM];
                load_face_pair<DIM>(wa_arr, f, w);
#pragma unroll
                for (int q = 0; q < NA; q++)
#pragma unroll
                    for (int d = 0; d < DIM; d++) gan[q][d] = w[1][d] * (ac[q] - an[q]);
            } else {
#pragma unroll
                for (int q = 0; q < NA; q++)
#pragma unroll
                    for (int d = 0; d < DIM; d++) gan[q][d] = gradA[(q * DIM + d) * N + n];
            }
            double vel[DIM], gT[DIM], gV[DIM][DIM], normal[DIM];
#pragma unroll
            for (int d = 0; d < DIM; d++) {
                vel[d] = 0.5 * ac[1 + d] + 0.5 * an[1 + d];
                gT[d] = 0.5 * gac[0][d] + 0.5 * gan[0][d];
                normal[d] = area * nhat[d];
#pragma unroll
                for (int q = 0; q < DIM; q++) gV[q][d] = 0.5 * gac[1 + q][d] + 0.5 * gan[1 + q][d];
            }
            diffusion_flux_add<double, DIM>(normal, ph.mu, ph.k, vel, gT, gV, flux);
            // species diffusion: -D rho_face sum_d n_d gradY_d, rho_face = 1/2 (rho_L + rho_R)
            cd = -ph.diff * (0.5 * uc[kRho] + 0.5 * un[kRho]) * 0.5;
        }
        const double sign = s == 0 ? -1.0 : 1.0;
#pragma unroll
        for (int p = 0; p < NE; p++) acc[p] += sign * flux[p];
        if (SPECIES) {
            sp_cl[nk] = sign * cl;
            sp_cr[nk] = sign * cr;
            sp_cd[nk] = sign * cd;
            sp_face[nk] = e;
            nk++;
        }
    }
    const double iv = m.inv_vol[c];
#pragma unroll
    for (int p = 0; p < NE; p++) {
        const double r = acc[p] * iv;
        R[p * N + c] = accumulate ? R[p * N + c] + r : r;
    }
    if (!SPECIES) return;
    // species pass: reconstructed densityYi on each side, upwinded by the stored coefficients
    const double rho_c = uc[kRho];
    for (int sp = 0; sp < ph.ns; sp++) {
        const int64_t comp = NE + sp;
        const double yc = U[comp * N + c];
        double gyc[DIM], gac_s[DIM];
#pragma unroll
        for (int d = 0; d < DIM; d++) {
            gyc[d] = RECON ? gradC[(comp * DIM + d) * N + c] : 0.0;
            gac_s[d] = VISC ? gradA[((DIM + 1 + sp) * DIM + d) * N + c] : 0.0;
        }
        double a = 0.0;
        for (int k = 0; k < nk; k++) {
            const int32_t e = sp_face[k];
            const int64_t f = e >> 1;
            const int s = e & 1;
            const int64_t n = m.fcell[(1 - s) * F + f];
            const bool ghost = n >= NI;
            const double yn = U[comp * N + n];
            double yself = yc, ynbr = yn;
            if (RECON) {
                double x[2][DIM];
                load_face_pair<DIM>(m.fdx, f, x);
                double ps = 0.0, pn = 0.0;
#pragma unroll
                for (int d = 0; d < DIM; d++) {
                    ps += gyc[d] * x[s][d];
                    if (!ghost) pn += gradC[(comp * DIM + d) * N + n] * x[1 - s][d];
                }
                yself += ps;
                ynbr += pn;
            }
            const double yL = s == 0 ? yself : ynbr;
            const double yR = s == 0 ? ynbr : yself;
            double contrib = sp_cl[k] * yL + sp_cr[k] * yR;
            if (VISC) {
                double gn[DIM];
                if (ghost) {
                    double w[2][DIM];
                    load_face_pair<DIM>(wa_arr, f, w);
                    const double rho_n = U[kRho * N + n];
#pragma unroll
                    for (int d = 0; d < DIM; d++) gn[d] = w[1][d] * (yc / rho_c - yn / rho_n);
                } else {
#pragma unroll
                    for (int d = 0; d < DIM; d++) gn[d] = gradA[((DIM + 1 + sp) * DIM + d) * N + n];
                }
                double nh[3], area;
                load_face_na(m.fna, f, nh, area);
                double dot = 0.0;
#pragma unroll
                for (int d = 0; d < DIM; d++) dot += area * nh[d] * (gac_s[d] + gn[d]);
                contrib += sp_cd[k] * dot;
            }
            a += contrib;
        }
        const double r = a * iv;
        R[comp * N + c] = accumulate ? R[comp * N + c] + r : r;
    }
}

// ---- compile-time variant selection
using GradKernel = void (*)(MeshView<double>, Phys, const double*, double*, double*);
using FluxKernel = void (*)(MeshView<double>, Phys, const double*, const double*, const double*, double*, int);

template <int DIM>
GradKernel pick_gradients(bool recon, bool visc, int lim) {
    if (!recon) return visc ? k_gradients<DIM, false, true, kLimNone> : nullptr;
    if (visc) {
        if (lim == kLimBJ) return k_gradients<DIM, true, true, kLimBJ>;
        if (lim == kLimPetsc) return k_gradients<DIM, true, true, kLimPetsc>;
        return k_gradients<DIM, true, true, kLimNone>;
    }
    if (lim == kLimBJ) return k_gradients<DIM, true, false, kLimBJ>;
    if (lim == kLimPetsc) return k_gradients<DIM, true, false, kLimPetsc>;
    return k_gradients<DIM, true, false, kLimNone>;
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

}  // namespace abl

struct abl_fv_solver {
    std::string last_error;
    abl_fv_physics_desc phys{};
    int form = 0;
    abl::Layout host;  // metadata only after upload (arrays released)
    abl::Phys ph{};
    abl::DevBuf<int32_t> ell, fcell, bface, bslot, slot_kind;
    abl::DevBuf<double> fna, fdx, fw, fwa, inv_vol, gradC, gradA, slot_state;
    abl::GradKernel grad_kernel = nullptr;
    abl::FluxKernel flux_kernel = nullptr;
    abl::DevBuf<unsigned long long> counter;
    std::vector<int32_t> h_slot_kind;
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
        v.fw = fw.ptr;
        v.fwa = fwa.ptr;
        v.inv_vol = inv_vol.ptr;
        return v;
    }
    int64_t device_bytes() const {
        return (int64_t)(ell.bytes() + fcell.bytes() + bface.bytes() + bslot.bytes() + slot_kind.bytes() + fna.bytes() + fdx.bytes() + fw.bytes() + fwa.bytes() + inv_vol.bytes() + gradC.bytes() +
                         gradA.bytes() + slot_state.bytes() + counter.bytes());
    }
};

namespace {

int fail(abl_fv_handle h, const std::string& msg) {
    if (h) h->last_error = msg;
    abl::set_global_error(msg);
    return 1;
}

inline unsigned grid_for(int64_t n) { return (unsigned)((n + abl::kBlock - 1) / abl::kBlock); }

}  // namespace

extern "C" const char* abl_fv_last_error(abl_fv_handle h) {
    if (h) return h->last_error.c_str();
    return abl::g_last_error.c_str();
}

extern "C" int abl_fv_create(const abl_fv_mesh_desc* mesh, const abl_fv_physics_desc* physics, int32_t form,
                             abl_fv_handle* out) {
    if (!mesh || !physics || !out) return fail(nullptr, "abl_fv_create: null argument");
    *out = nullptr;
    if (form != ABL_FV_FORM_FP64) return fail(nullptr, "abl_fv_create: only the FP64 form is available in this build");
    const abl_fv_physics_desc& p = *physics;
    if (!(p.gamma > 1.0) || !(p.Rgas > 0.0)) return fail(nullptr, "abl_fv_create: PerfectGas needs gamma > 1 and Rgas > 0");
    if (p.flux_calculator < ABL_FV_FLUX_NONE || p.flux_calculator > ABL_FV_FLUX_RUSANOV)
        return fail(nullptr, "abl_fv_create: unknown flux calculator");
    if (p.reconstruction != ABL_FV_RECON_NONE && p.reconstruction != ABL_FV_RECON_LEASTSQUARES)
        return fail(nullptr, "abl_fv_create: unknown reconstruction");
    if (p.limiter < ABL_FV_LIMITER_NONE || p.limiter > ABL_FV_LIMITER_BARTH_JESPERSEN)
        return fail(nullptr, "abl_fv_create: unknown limiter");
    std::unique_ptr<abl_fv_solver> s(new abl_fv_solver());
    try {
        s->phys = p;
        s->form = form;
        abl::build_layout(*mesh, p, s->host);
        abl::Layout& L = s->host;
        s->recon = p.reconstruction == ABL_FV_RECON_LEASTSQUARES && p.flux_calculator != ABL_FV_FLUX_NONE;
        s->visc = p.k != 0.0 || p.mu != 0.0 || p.diff != 0.0;
        abl::Phys& ph = s->ph;
        ph.gamma = p.gamma;
        ph.R = p.Rgas;
        ph.cv = p.Rgas / (p.gamma - 1.0);
        ph.icv = 1.0 / ph.cv;
        ph.mu = p.mu;
        ph.k = p.k;
        ph.diff = p.diff;
        ph.max_lim_grad = p.max_lim_grad;
        ph.flux = p.flux_calculator;
        ph.recon = s->recon;
        ph.limiter = p.limiter;
        ph.ns = p.n_species;
        ph.visc = s->visc;
        ph.sdiff = s->visc && p.n_species > 0;
        const int lim = p.limiter == ABL_FV_LIMITER_NONE ? abl::kLimNone
                        : p.limiter == ABL_FV_LIMITER_BARTH_JESPERSEN ? abl::kLimBJ : abl::kLimPetsc;
        const bool species = p.n_species > 0;
        if (mesh->dim == 3) {
            s->grad_kernel = abl::pick_gradients<3>(s->recon, s->visc, lim);
            s->flux_kernel = abl::pick_flux<3>(s->recon, s->visc, species);
        } else {
            s->grad_kernel = abl::pick_gradients<2>(s->recon, s->visc, lim);
            s->flux_kernel = abl::pick_flux<2>(s->recon, s->visc, species);
        }

        s->ell.upload(L.ell);
        s->fcell.upload(L.fcell);
        s->fna.upload(L.fna);
        s->fdx.upload(L.fdx);
        s->fw.upload(L.fw);
        s->fwa.upload(L.fwa);
        s->inv_vol.upload(L.inv_vol);
        const int64_t N = L.N;
        if (s->recon) {
            s->gradC.alloc((size_t)L.nc * L.dim * N);
            HIP_CHECK(hipMemset(s->gradC.ptr, 0, s->gradC.bytes()));  // ghost rows stay zero
        }
        if (s->visc) {
            s->gradA.alloc((size_t)(1 + L.dim + L.ns) * L.dim * N);
            HIP_CHECK(hipMemset(s->gradA.ptr, 0, s->gradA.bytes()));
        }
        std::vector<int32_t> bf, bs;
        for (auto& b : L.bfaces) {
            bf.push_back(b.face);
            bs.push_back(b.slot);
        }
        s->n_bfaces = (int64_t)bf.size();
        s->bface.upload(bf);
        s->bslot.upload(bs);
        s->h_slot_kind.assign(abl::kMaxBcSlots, ABL_FV_BC_NONE);
        s->h_slot_state.assign((size_t)abl::kMaxBcSlots * L.nc, 0.0);
        s->h_slot_theta.assign(abl::kMaxBcSlots, {0.0, 0.0, 0.0});
        s->slot_kind.upload(s->h_slot_kind);
        s->slot_state.upload(s->h_slot_state);
        s->counter.alloc(1);
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
    if (bc->kind != ABL_FV_BC_INFLOW && bc->kind != ABL_FV_BC_OUTFLOW && bc->kind != ABL_FV_BC_NONE)
        return fail(h, "abl_fv_set_boundary: unknown boundary kind");
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
    try {
        HIP_CHECK(hipMemcpy(h->slot_kind.ptr, h->h_slot_kind.data(), h->slot_kind.bytes(), hipMemcpyHostToDevice));
        HIP_CHECK(hipMemcpy(h->slot_state.ptr, h->h_slot_state.data(), h->slot_state.bytes(), hipMemcpyHostToDevice));
    } catch (const std::exception& e) {
        return fail(h, e.what());
    }
    return 0;
}

extern "C" int abl_fv_rhs(abl_fv_handle h, double time, double* U, double* R, int32_t accumulate, abl_fv_stream stream) {
    (void)time;
    if (!h || !U || !R) return fail(h, "abl_fv_rhs: null argument");
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
        const auto v = h->view();
        const dim3 grid(grid_for(L.NI)), block(abl::kBlock);
        if (h->recon || h->visc) {
            h->timed(ABL_FV_PHASE_GRADIENTS, st, [&] {
                hipLaunchKernelGGL(h->grad_kernel, grid, block, 0, st, v, h->ph, U, h->gradC.ptr, h->gradA.ptr);
            });
            if (L.ns)
                h->timed(ABL_FV_PHASE_SPECIES_GRADIENTS, st, [&] {
                    if (L.dim == 3)
                        hipLaunchKernelGGL(abl::k_species_gradients<3>, grid, block, 0, st, v, h->ph, U, h->gradC.ptr,
                                           h->gradA.ptr);
                    else
                        hipLaunchKernelGGL(abl::k_species_gradients<2>, grid, block, 0, st, v, h->ph, U, h->gradC.ptr,
                                           h->gradA.ptr);
                });
        }
        h->timed(ABL_FV_PHASE_FLUX, st, [&] {
            hipLaunchKernelGGL(h->flux_kernel, grid, block, 0, st, v, h->ph, U, h->gradC.ptr, h->gradA.ptr, R, accumulate);
        });
        if (!accumulate && L.N > L.NI)
            hipLaunchKernelGGL(abl::k_zero_rows, dim3(grid_for((L.N - L.NI) * L.nc)), block, 0, st, R, L.N, L.NI, L.nc);
        HIP_CHECK(hipGetLastError());
    } catch (const std::exception& e) {
        return fail(h, std::string("abl_fv_rhs: ") + e.what());
    }
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

extern "C" int abl_fv_rhs_vjp(abl_fv_handle h, double, double*, const double*, double*, double*, abl_fv_stream) {
    return fail(h, "abl_fv_rhs_vjp: not available in this build");
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
