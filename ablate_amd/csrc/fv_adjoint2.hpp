// Reverse-mode pieces for the second-order adjoint: transposes of the Navier-Stokes viscous face flux, of the
// stress / heat-flux map and of the aux (temperature, velocity) and mass-fraction decodes.  Each adds into its
// output arrays; the forward definitions they reverse are NavierStokesTransport::DiffusionFlux and
// CompressibleFlowComputeStressTensor (navierStokesTransport.cpp:454-532), the PerfectGas temperature
// (perfectGas.cpp:76-86) and SpeciesTransport::UpdateAuxMassFractionField (speciesTransport.cpp:98-110).
#pragma once
#include "fv_device.hpp"

namespace abl {

// Forward: flux[rhoU_c] += -sum_d normal_d tq[sym(c,d)];
//          flux[rhoE]   += -sum_d normal_d (sum_c vel_c tq[sym(d,c)] + tq[NTAU + d]).
template <int DIM>
__device__ inline void viscous_flux_vjp(const double* normal, const double* vel, const double* tq, const double* fb,
                                        double* tqb, double* velb) {
    constexpr int NTAU = ViscTerms<DIM>::NTAU;
#pragma unroll
    for (int c = 0; c < DIM; c++)
#pragma unroll
        for (int d = 0; d < DIM; d++) tqb[sym_index(DIM, c, d)] -= normal[d] * fb[kRhoU + c];
#pragma unroll
    for (int d = 0; d < DIM; d++) {
#pragma unroll
        for (int c = 0; c < DIM; c++) {
            tqb[sym_index(DIM, d, c)] -= normal[d] * vel[c] * fb[kRhoE];
            velb[c] -= normal[d] * tq[sym_index(DIM, d, c)] * fb[kRhoE];
        }
        tqb[NTAU + d] -= normal[d] * fb[kRhoE];
    }
}

// Transpose of stress_heat: tq = S(gA) with gA[0][d] = dT/dx_d, gA[1 + c][d] = du_c/dx_d.
template <int DIM>
__device__ inline void stress_heat_vjp(double mu, double k, const double* tqb, double (*gAb)[DIM]) {
    constexpr int NTAU = ViscTerms<DIM>::NTAU;
    double divb = 0.0;
#pragma unroll
    for (int c = 0; c < DIM; c++)
#pragma unroll
        for (int d = c; d < DIM; d++) {
            const double t = tqb[sym_index(DIM, c, d)];
            if (c == d) {
                gAb[1 + c][c] += 2.0 * mu * t;
                divb -= 2.0 * mu * t / 3.0;
            } else {
                gAb[1 + c][d] += mu * t;
                gAb[1 + d][c] += mu * t;
            }
        }
#pragma unroll
    for (int c = 0; c < DIM; c++) gAb[1 + c][c] += divb;
#pragma unroll
    for (int d = 0; d < DIM; d++) gAb[0][d] += k * tqb[NTAU + d];
}

// Adjoint of the aux decode a = (T, vel) of a cell state u: T = (rhoE/rho - |v|^2/2) / cv, v = rhoU / rho.
template <int DIM>
__device__ inline void aux_vjp(const double* u, const Phys& ph, const double* ab, double* ub) {
    const double irho = 1.0 / u[kRho];
    const double Tb = ab[0];
    ub[kRhoE] += Tb * ph.icv * irho;
    ub[kRho] -= Tb * ph.icv * u[kRhoE] * irho * irho;
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        const double v = u[kRhoU + d] * irho;
        const double vb = ab[1 + d] - Tb * v * ph.icv;
        ub[kRhoU + d] += vb * irho;
        ub[kRho] -= vb * v * irho;
    }
}

// Adjoint of Y = rhoY / rho: accumulates into (rhoY_b, rho_b).
__device__ inline void massfrac_vjp(double rho, double rhoY, double Yb, double& rhoYb, double& rhob) {
    rhoYb += Yb / rho;
    rhob -= Yb * rhoY / (rho * rho);
}

}  // namespace abl
