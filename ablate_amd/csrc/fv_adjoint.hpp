// Hand-written reverse mode (vector-Jacobian product) of the first-order face fluxes: the adjoint of
// NavierStokesTransport::AdvectionFlux + SpeciesTransport::AdvectionFlux with the AUSM calculator
// (navierStokesTransport.cpp:99-207, speciesTransport.cpp:263-341, ausm.cpp:2-43) and of the build-defined
// Rusanov flux.  The forward intermediates are recomputed, never stored.
#pragma once
#include "fv_device.hpp"

namespace abl {

// Forward decode of one side, keeping what the reverse sweep needs (PerfectGas T-route, perfectGas.cpp:67-174).
template <int DIM>
struct SideFwd {
    double rho, irho, vel[DIM], eint, T, p, a, un, H;
};

template <int DIM>
__device__ inline void side_forward(const double* u, const double* nhat, const Phys& ph, SideFwd<DIM>& s) {
    s.rho = u[kRho];
    s.irho = 1.0 / s.rho;
    double ke = 0.0;
    s.un = 0.0;
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        s.vel[d] = u[kRhoU + d] * s.irho;
        ke += s.vel[d] * s.vel[d];
        s.un += s.vel[d] * nhat[d];
    }
    s.eint = u[kRhoE] * s.irho - 0.5 * ke;
    s.T = s.eint * ph.icv;
    s.p = ph.R * s.rho * s.T;
    s.a = sqrt(ph.gamma * s.p * s.irho);
    double v2 = 0.0;
#pragma unroll
    for (int d = 0; d < DIM; d++) v2 += s.vel[d] * s.vel[d];
    s.H = s.T * ph.cv + 0.5 * v2 + s.p * s.irho;
}

// Adjoint of side_forward: given the adjoints of (rho, vel, p, a, un, H) (rho_b collects direct uses of rho),
// accumulates the adjoint of the conserved state into ub[0..DIM+1].
template <int DIM>
__device__ inline void side_reverse(const SideFwd<DIM>& s, const double* nhat, const Phys& ph, double rho_b,
                                    double* vel_b, double p_b, double a_b, double un_b, double H_b, double* ub) {
    // H = T cv + |v|^2 / 2 + p / rho
    double T_b = H_b * ph.cv;
#pragma unroll
    for (int d = 0; d < DIM; d++) vel_b[d] += H_b * s.vel[d];
    p_b += H_b * s.irho;
    rho_b -= H_b * s.p * s.irho * s.irho;
    // un = vel . nhat
#pragma unroll
    for (int d = 0; d < DIM; d++) vel_b[d] += un_b * nhat[d];
    // a = sqrt(gamma p / rho)
    if (a_b != 0.0) {
        const double t = a_b * 0.5 * ph.gamma / (s.a * s.rho);
        p_b += t;
        rho_b -= t * s.p * s.irho;
    }
    // p = R rho T
    rho_b += p_b * ph.R * s.T;
    T_b += p_b * ph.R * s.rho;
    // T = eint / cv, eint = rhoE / rho - |v|^2 / 2
    // eint = rhoE / rho - ke with ke = |v|^2 / 2: d/drhoE = 1/rho, d/drho = -(rhoE/rho)/rho = -(eint + ke)/rho,
    // d/dv = -v
    const double eint_b = T_b * ph.icv;
    double ke = 0.0;
#pragma unroll
    for (int d = 0; d < DIM; d++) ke += 0.5 * s.vel[d] * s.vel[d];
    ub[kRhoE] += eint_b * s.irho;
    rho_b -= eint_b * (s.eint + ke) * s.irho;
#pragma unroll
    for (int d = 0; d < DIM; d++) vel_b[d] -= eint_b * s.vel[d];
    // vel = rhoU / rho
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        ub[kRhoU + d] += vel_b[d] * s.irho;
        rho_b -= vel_b[d] * s.vel[d] * s.irho;
    }
    ub[kRho] += rho_b;
}

// Split-Mach derivatives of AUSM (ausm.cpp:4-22): value and d/dM of (M-, P-) for the right state and (M+, P+) for
// the left state.
__device__ inline void ausm_minus(double M, double& sM, double& sP, double& dsM, double& dsP) {
    if (fabs(M) <= 1.0) {
        sM = -0.25 * (M - 1) * (M - 1);
        dsM = -0.5 * (M - 1);
        sP = -sM * (2 + M);
        dsP = -dsM * (2 + M) - sM;
    } else {
        sM = 0.5 * (M - fabs(M));
        dsM = M < 0 ? 1.0 : 0.0;
        sP = sM / M;
        dsP = 0.0;  // sP is 0 or 1 on the supersonic branches
    }
}
__device__ inline void ausm_plus(double M, double& sM, double& sP, double& dsM, double& dsP) {
    if (fabs(M) <= 1.0) {
        sM = 0.25 * (M + 1) * (M + 1);
        dsM = 0.5 * (M + 1);
        sP = sM * (2 - M);
        dsP = dsM * (2 - M) - sM;
    } else {
        sM = 0.5 * (M + fabs(M));
        dsM = M > 0 ? 1.0 : 0.0;
        sP = sM / M;
        dsP = 0.0;
    }
}

// VJP of the AUSM face flux.  uL/uR: euler states; fb: adjoint of the euler face flux [NE]; species enter through
// QL/QR = sum_sp fb_sp rhoY_sp of each side (fb_sp: adjoint of the species face flux).  Accumulates the euler adjoints
// of both sides into ubL/ubR and returns the per-species multipliers: rhoY_bar_sp(side) += coef(side) * fb_sp.
template <int DIM>
__device__ inline void ausm_flux_vjp(const double* uL, const double* uR, double QL, double QR, const double* nhat,
                                     double area, const Phys& ph, const double* fb, double* ubL, double* ubR,
                                     double& coefL, double& coefR) {
    SideFwd<DIM> L, R;
    side_forward<DIM>(uL, nhat, ph, L);
    side_forward<DIM>(uR, nhat, ph, R);
    const double Mm = R.un / R.a, Mp = L.un / L.a;
    double sMm, sPm, dsMm, dsPm, sMp, sPp, dsMp, dsPp;
    ausm_minus(Mm, sMm, sPm, dsMm, dsPm);
    ausm_plus(Mp, sMp, sPp, dsMp, dsPp);
    const double m = sMm + sMp;
    const bool right = m < 0;  // SpeciesTransport takes the right state for any non-LEFT direction
    const SideFwd<DIM>& S = right ? R : L;
    const double mdot = m * S.a * S.rho;
    // reverse of F_rho = mdot A; F_E = H_s mdot A; F_u = v_s mdot A + p12 n A; F_sp = mdot A rhoY_s / rho_s
    const double QS = right ? QR : QL;
    double mdot_b = area * fb[kRho] + area * S.H * fb[kRhoE] + area * S.irho * QS;
    const double H_b = area * mdot * fb[kRhoE];
    double velS_b[DIM], p12_b = 0.0;
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        mdot_b += area * S.vel[d] * fb[kRhoU + d];
        velS_b[d] = area * mdot * fb[kRhoU + d];
        p12_b += area * nhat[d] * fb[kRhoU + d];
    }
    const double coefS = area * mdot * S.irho;
    coefL = right ? 0.0 : coefS;
    coefR = right ? coefS : 0.0;
    // mdot = m a_s rho_s
    const double m_b = mdot_b * S.a * S.rho;
    const double aS_b = mdot_b * m * S.rho;
    const double rhoS_b = mdot_b * m * S.a - area * mdot * S.irho * S.irho * QS;
    // p12 = pR sPm + pL sPp
    const double pR_b = p12_b * sPm, pL_b = p12_b * sPp;
    const double sPm_b = p12_b * R.p, sPp_b = p12_b * L.p;
    // m = sMm + sMp; Mm = uR_n / aR, Mp = uL_n / aL
    const double Mm_b = m_b * dsMm + sPm_b * dsPm;
    const double Mp_b = m_b * dsMp + sPp_b * dsPp;
    const double unR_b = Mm_b / R.a, aR_b = -Mm_b * Mm / R.a;
    const double unL_b = Mp_b / L.a, aL_b = -Mp_b * Mp / L.a;
    double velL_b[DIM], velR_b[DIM];
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        velL_b[d] = right ? 0.0 : velS_b[d];
        velR_b[d] = right ? velS_b[d] : 0.0;
    }
    side_reverse<DIM>(L, nhat, ph, right ? 0.0 : rhoS_b, velL_b, pL_b, aL_b + (right ? 0.0 : aS_b), unL_b,
                      right ? 0.0 : H_b, ubL);
    side_reverse<DIM>(R, nhat, ph, right ? rhoS_b : 0.0, velR_b, pR_b, aR_b + (right ? aS_b : 0.0), unR_b,
                      right ? H_b : 0.0, ubR);
}

// VJP of the Rusanov full-vector flux F = A [1/2 (f_L + f_R) - 1/2 lambda (U_R - U_L)]; species as in
// ausm_flux_vjp (QL, QR, per-species multipliers coefL/coefR).
template <int DIM>
__device__ inline void rusanov_flux_vjp(const double* uL, const double* uR, double QL, double QR, double QdL,
                                        const double* nhat, double area, const Phys& ph, const double* fb, double* ubL,
                                        double* ubR, double& coefL, double& coefR) {
    SideFwd<DIM> L, R;
    side_forward<DIM>(uL, nhat, ph, L);
    side_forward<DIM>(uR, nhat, ph, R);
    const double sL = fabs(L.un) + L.a, sR = fabs(R.un) + R.a;
    const bool left_max = !(sL < sR);  // pmax(sL, sR) picks sL unless sL < sR
    const double lam = left_max ? sL : sR;
    constexpr int NE = DIM + 2;
    double lam_b = 0.0, unL_b = 0.0, unR_b = 0.0, pL_b = 0.0, pR_b = 0.0;
#pragma unroll
    for (int c = 0; c < NE; c++) {
        const double g = area * fb[c];
        const double gh = 0.5 * g;
        lam_b -= gh * (uR[c] - uL[c]);
        ubR[c] -= gh * lam;
        ubL[c] += gh * lam;
        ubL[c] += gh * L.un;
        ubR[c] += gh * R.un;
        if (c == kRho) {
            unL_b += gh * uL[c];
            unR_b += gh * uR[c];
        } else if (c == kRhoE) {
            pL_b += gh * L.un;
            pR_b += gh * R.un;
            unL_b += gh * (uL[c] + L.p);
            unR_b += gh * (uR[c] + R.p);
        } else {
            unL_b += gh * uL[c];
            unR_b += gh * uR[c];
            pL_b += gh * nhat[c - kRhoU];
            pR_b += gh * nhat[c - kRhoU];
        }
    }
    // species: F_sp = A [1/2 (Y_L un_L + Y_R un_R) - 1/2 lambda (Y_R - Y_L)] with Y = rhoY
    unL_b += 0.5 * area * QL;
    unR_b += 0.5 * area * QR;
    lam_b -= 0.5 * area * QdL;  // QdL = sum_sp fb_sp (rhoY_R - rhoY_L)
    coefL = area * (0.5 * L.un + 0.5 * lam);
    coefR = area * (0.5 * R.un - 0.5 * lam);
    // d|u_n|/du_n = sign(u_n) with sign(0) = 0 (the central-difference value at the kink)
    double aL_b = 0.0, aR_b = 0.0;
    if (left_max) {
        unL_b += lam_b * (double)((L.un > 0) - (L.un < 0));
        aL_b += lam_b;
    } else {
        unR_b += lam_b * (double)((R.un > 0) - (R.un < 0));
        aR_b += lam_b;
    }
    double velL_b[DIM], velR_b[DIM];
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        velL_b[d] = 0.0;
        velR_b[d] = 0.0;
    }
    side_reverse<DIM>(L, nhat, ph, 0.0, velL_b, pL_b, aL_b, unL_b, 0.0, ubL);
    side_reverse<DIM>(R, nhat, ph, 0.0, velR_b, pR_b, aR_b, unR_b, 0.0, ubR);
}

}  // namespace abl
