// Device-side building blocks of the FV RHS: thermodynamics (PerfectGas), flux calculators, the Navier-Stokes
// advection / diffusion fluxes and the gradient limiters.  Templated on the arithmetic type so that the FP64
// and the mixed-precision forms share one source.
#pragma once
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>

namespace abl {

enum : int { kRho = 0, kRhoE = 1, kRhoU = 2 };

// Device code of an INFLOW slot with enforceAtFace (essentialGhost.cpp:15-21: ghost = 2 state - interior); the other
// slot kinds keep their ABL_FV_BC_* values on the device.
constexpr int kBcInflowAtFace = 16;

// Physics constants passed by value to every kernel.
struct Phys {
    double gamma, R, cv, icv, mu, k, diff, max_lim_grad;
    double minf;       // AusmpUp reference Mach number (0: fa = 1)
    double pgs_alpha;  // pressure-gradient-scaling factor seen by AusmpUp (1 without PGS)
    int compact;       // CompactCompressibleNSSpeciesTransport: NA directions average the species
    int flux, recon, limiter, ns;
    int visc;   // Navier-Stokes diffusion flux active (k or mu nonzero, or any transport active)
    int sdiff;  // species diffusion active
    // Sutherland transport: the cell records hold the strain-rate and temperature-gradient terms with mu = k = 1
    // (mu, k, diff above are then 1 or 0 per enabled property) and every face scales them by mu(T_face)
    int suth;
    double cp;  // specific heat at constant pressure (PerfectGas: gamma R / (gamma - 1))
};

// Sutherland's law with the reference's constants (eos/transport/sutherland.hpp: mu_0 = 1.716e-5, T_0 = 273,
// S = 111, Pr = Sc = 0.707; sutherland.cpp:17-19, 39-40): mu(T) = mu_0 (T / T_0)^(3/2) (T_0 + S) / (T + S).
constexpr double kSuthMu0 = 1.716e-5, kSuthT0 = 273.0, kSuthS = 111.0, kSuthPr = 0.707, kSuthSc = 0.707;
__device__ __host__ inline double sutherland_mu(double T) {
    const double r = T / kSuthT0;
    return kSuthMu0 * sqrt(r) * r * (kSuthT0 + kSuthS) / (T + kSuthS);
}

// Geometry view (G = storage type of the face geometry: double for FP64, float for the mixed form).
template <typename G>
struct MeshView {
    int64_t N, NI, F;
    int K;
    const int32_t* __restrict__ ell;
    const int32_t* __restrict__ fcell;
    const G* __restrict__ fna;  // [F][4]: unit normal, area
    const G* __restrict__ fdx;  // [F][2][DIM]: x_face - x_cell per side
    const G* __restrict__ fw;   // [F][2][DIM]: cell least-squares weights per side (unused by the kernels)
    const G* __restrict__ fwa;  // [F][2][DIM]: aux (face-stencil) weights, or null when there are no boundaries
    const double* __restrict__ inv_vol;
    // Slot-major copies of the per-side geometry, entry (k, d, c) at (k * DIM + d) * NI + c, for KM = slot width
    // slots (padding slots: neighbour -1, zero weights).  A lane per cell streams its own records.
    const int32_t* __restrict__ snbr;  // [KM][NI] cell across slot k, -1 = no face
    const G* __restrict__ sw;          // [KM][DIM][NI] cell least-squares weight of the slot's side
    const G* __restrict__ swa;         // [KM][DIM][NI] aux weights, or null = sw
    const G* __restrict__ slim;        // [KM][DIM][NI] limiter geometry: x_f - x_c (BJ) or x_n - x_c (PETSc family)
};

// Block index remapped so that each XCD (workgroups are dispatched round-robin over 8 XCDs) walks a contiguous
// range of cells, keeping neighbour reuse inside one XCD's L2.
__device__ inline int64_t xcd_block_index() {
    const uint32_t nb = gridDim.x, b = blockIdx.x;
    const uint32_t q = nb >> 3, r = nb & 7u, x = b & 7u;
    return (int64_t)(x * q + (x < r ? x : r) + (b >> 3));
}
__device__ inline int64_t xcd_cell_index() { return xcd_block_index() * blockDim.x + threadIdx.x; }

// Face records are read with vector loads (16 bytes for FP64 records, 8/16 bytes for FP32 ones).
__device__ inline void load_face_na(const double* __restrict__ fna, int64_t f, double* nhat, double& area) {
    const double2* p = reinterpret_cast<const double2*>(fna + 4 * f);
    const double2 a = p[0], b = p[1];
    nhat[0] = a.x;
    nhat[1] = a.y;
    nhat[2] = b.x;
    area = b.y;
}
__device__ inline void load_face_na(const float* __restrict__ fna, int64_t f, float* nhat, float& area) {
    const float4 a = *reinterpret_cast<const float4*>(fna + 4 * f);
    nhat[0] = a.x;
    nhat[1] = a.y;
    nhat[2] = a.z;
    area = a.w;
}
template <typename G>
struct Vec2T;
template <>
struct Vec2T<double> {
    using type = double2;
};
template <>
struct Vec2T<float> {
    using type = float2;
};
// Loads the per-face pair record x[2][DIM] (both sides) of fdx / fw / fwa.
template <int DIM, typename G>
__device__ inline void load_face_pair(const G* __restrict__ arr, int64_t f, G (*x)[DIM]) {
    using V2 = typename Vec2T<G>::type;
    const V2* p = reinterpret_cast<const V2*>(arr + 2 * DIM * f);
    G t[2 * DIM];
#pragma unroll
    for (int i = 0; i < DIM; i++) {
        const V2 v = p[i];
        t[2 * i] = v.x;
        t[2 * i + 1] = v.y;
    }
#pragma unroll
    for (int s = 0; s < 2; s++)
#pragma unroll
        for (int d = 0; d < DIM; d++) x[s][d] = t[s * DIM + d];
}

// Face slot k of interior cell c in the cell -> face table.  Loops over slots are unrolled to a compile-time
// width so that every slot's loads issue up front; padding slots map to face 0 / the cell itself (valid addresses)
// and are masked out of the arithmetic.
struct Slot {
    bool valid;
    int s;
    int64_t f, n;
};
template <typename MV>
__device__ inline Slot face_slot(const MV& m, int k, int64_t c) {
    const int32_t e = k < m.K ? m.ell[k * m.NI + c] : -1;
    Slot sl;
    sl.valid = e >= 0;
    sl.f = sl.valid ? (int64_t)(e >> 1) : 0;
    sl.s = sl.valid ? (e & 1) : 0;
    sl.n = sl.valid ? (int64_t)m.fcell[(1 - sl.s) * m.F + sl.f] : c;
    return sl;
}

// PETSc's PetscMin/PetscMax: the comparison order decides NaN propagation.
template <typename T>
__device__ inline T pmin(T a, T b) { return a < b ? a : b; }
template <typename T>
__device__ inline T pmax(T a, T b) { return a < b ? b : a; }

// Flux-calculator direction (fluxCalculator.hpp:12)
enum : int { kNA = 0, kLeft = 1, kRight = 2 };

// Liou-Steffen AUSM (ausm.cpp:2-43).
template <typename T>
__device__ inline int ausm(T uL, T aL, T rhoL, T pL, T uR, T aR, T rhoR, T pR, T& massFlux, T& p12) {
    const T Mm = uR / aR;
    T sMm, sPm;
    if (fabs(Mm) <= T(1)) {
        sMm = T(-0.25) * (Mm - T(1)) * (Mm - T(1));
        sPm = -sMm * (T(2) + Mm);
    } else {
        sMm = T(0.5) * (Mm - fabs(Mm));
        sPm = sMm / Mm;
    }
    const T Mp = uL / aL;
    T sMp, sPp;
    if (fabs(Mp) <= T(1)) {
        sMp = T(0.25) * (Mp + T(1)) * (Mp + T(1));
        sPp = sMp * (T(2) - Mp);
    } else {
        sMp = T(0.5) * (Mp + fabs(Mp));
        sPp = sMp / Mp;
    }
    const T m = sMm + sMp;
    int dir;
    if (m < T(0)) {
        massFlux = m * aR * rhoR;
        dir = kRight;
    } else {
        massFlux = m * aL * rhoL;
        dir = kLeft;
    }
    p12 = pR * sPm + pL * sPp;
    return dir;
}

// Liou's AUSM+-up (J. Comput. Phys. 214 (2006) 137-170) as fluxCalculator::AusmpUp (ausmpUp.cpp:5-97; constants
// beta 1/8, Kp 1/4, Ku 3/4, sigma 1/4 from ausmpUp.hpp:21-24).  s = +1 / -1 selects the split polynomial:
// M1 = 1/2 (m + s|m|), M2 = s/4 (m + s)^2, M4 = M1 (|m| >= 1) or M2 (1 - 16 s beta M2(-s)),
// P5 = M1 / (m + 1e-30) (|m| >= 1) or M2 ((2s - m) - 16 s a m M2(-s)), a = 3/16 (5 fa^2 - 4).
template <typename T>
__device__ inline T aup_m1(T m, T s) { return T(0.5) * (m + s * fabs(m)); }
template <typename T>
__device__ inline T aup_m2(T m, T s) {
    const T q = m + s;
    return s * T(0.25) * (q * q);
}
template <typename T>
__device__ inline T aup_m4(T m, T s) {
    if (fabs(m) >= T(1)) return aup_m1(m, s);
    return aup_m2(m, s) * (T(1) - s * T(2) * aup_m2(m, -s));  // 16 beta = 2
}
template <typename T>
__device__ inline T aup_p5(T m, T s, T fa) {
    if (fabs(m) >= T(1)) return aup_m1(m, s) / (m + T(1e-30));
    const T a = T(3.0 / 16.0) * (T(-4) + T(5) * fa * fa);
    return aup_m2(m, s) * ((s * T(2) - m) - s * T(16) * a * m * aup_m2(m, -s));
}
template <typename T>
__device__ inline int ausmpup(T minf, T alpha, T uL, T aL, T rhoL, T pL, T uR, T aR, T rhoR, T pR, T& massFlux,
                              T& p12) {
    const T rho12 = T(0.5) * (rhoL + rhoR);
    const T a12 = T(0.5) * (aL + aR) / alpha;
    const T mL = uL / a12, mR = uR / a12;
    const T mBar2 = (uL * uL + uR * uR) / (T(2) * a12 * a12);
    T fa = T(1);
    if (minf > T(0)) {
        const T mO = sqrt(pmin(T(1), pmax(mBar2, minf * minf)));
        fa = mO * (T(2) - mO);
    }
    // products associate left to right as in ausmpUp.cpp:38,53
    const T m12 = aup_m4(mL, T(1)) + aup_m4(mR, T(-1)) -
                  (T(0.25) / fa) * pmax(T(1) - T(0.25) * mBar2, T(0)) * (pR - pL) / (rho12 * a12 * a12 * alpha * alpha);
    int dir;
    if (m12 > T(0)) {
        dir = kLeft;
        massFlux = a12 * m12 * rhoL;
    } else {
        dir = kRight;
        massFlux = a12 * m12 * rhoR;
    }
    const T pp = aup_p5(mL, T(1), fa), pm = aup_p5(mR, T(-1), fa);
    p12 = (pp * pL + pm * pR - T(0.75) * pp * pm * rho12 * fa * a12 * a12 * alpha * alpha * (mR - mL)) / (alpha * alpha);
    return dir;
}

// Flux calculators behind FluxCalculatorFunction: Ausm, AverageFlux (averageFlux.cpp:2-10), OffFlux, AusmpUp.
// AUP = false compiles AusmpUp out: its temporaries would otherwise raise the register count of kernels that never
// run it (the mixed form's FP64 route: 156 -> 178 VGPRs, one wave per SIMD less).
template <typename T, bool AUP = true>
__device__ inline int flux_calculator(int type, const Phys& ph, T uL, T aL, T rhoL, T pL, T uR, T aR, T rhoR, T pR,
                                      T& mdot, T& p12) {
    if (type == 1) return ausm(uL, aL, rhoL, pL, uR, aR, rhoR, pR, mdot, p12);
    if (AUP && type == 5)
        return ausmpup(T(ph.minf), T(ph.pgs_alpha), uL, aL, rhoL, pL, uR, aR, rhoR, pR, mdot, p12);
    if (type == 2) {
        mdot = T(0.5) * (uL * rhoL + uR * rhoR);
        p12 = T(0.5) * (pL + pR);
        return kNA;
    }
    mdot = T(0);
    p12 = T(0);
    return kNA;
}

// Decoded face-side state (PerfectGas T-route of AdvectionFlux, navierStokesTransport.cpp:122-140 and
// perfectGas.cpp:67-113,165-174).
template <typename T, int DIM>
struct SideState {
    T rho, irho, vel[DIM], un, e, a, p;
};

template <typename T, int DIM>
__device__ inline void decode_side(const T* u, const T* nhat, const Phys& ph, SideState<T, DIM>& s) {
    s.rho = u[kRho];
    s.irho = T(1) / s.rho;
    T ke = T(0);
    s.un = T(0);
    for (int d = 0; d < DIM; d++) {
        s.vel[d] = u[kRhoU + d] * s.irho;
        ke += s.vel[d] * s.vel[d];
        s.un += s.vel[d] * nhat[d];
    }
    ke *= T(0.5);
    const T eint = u[kRhoE] * s.irho - ke;
    const T Tmp = eint * T(ph.icv);
    s.e = Tmp * T(ph.cv);
    s.p = T(ph.R) * s.rho * Tmp;
    s.a = sqrt(T(ph.gamma) * s.p * s.irho);
}

// NavierStokesTransport::AdvectionFlux (navierStokesTransport.cpp:99-207) for euler components, plus the mass
// flux per area and the calculator's direction (kLeft / kRight / kNA) for the species fluxes.  flux[] receives
// dim+2 values.
template <typename T, int DIM, bool AUP = true>
__device__ inline void advection_flux(int type, const T* uL, const T* uR, const T* nhat, T area, const Phys& ph, T* flux,
                                      T& mdotA, int& direction, T& rhoL_out, T& rhoR_out) {
    SideState<T, DIM> L, R;
    decode_side<T, DIM>(uL, nhat, ph, L);
    decode_side<T, DIM>(uR, nhat, ph, R);
    T massFlux, p12;
    const int dir = flux_calculator<T, AUP>(type, ph, L.un, L.a, L.rho, L.p, R.un, R.a, R.rho, R.p, massFlux, p12);
    const T mA = massFlux * area;
    const T p12A = p12 * area;
    flux[kRho] = mA;
    if (dir == kLeft) {
        T v2 = T(0);
        for (int d = 0; d < DIM; d++) v2 += L.vel[d] * L.vel[d];
        const T H = L.e + v2 / T(2) + L.p * L.irho;
        flux[kRhoE] = H * mA;
        for (int d = 0; d < DIM; d++) flux[kRhoU + d] = L.vel[d] * mA + p12A * nhat[d];
    } else if (dir == kRight) {
        T v2 = T(0);
        for (int d = 0; d < DIM; d++) v2 += R.vel[d] * R.vel[d];
        const T H = R.e + v2 / T(2) + R.p * R.irho;
        flux[kRhoE] = H * mA;
        for (int d = 0; d < DIM; d++) flux[kRhoU + d] = R.vel[d] * mA + p12A * nhat[d];
    } else {
        T vl2 = T(0), vr2 = T(0);
        for (int d = 0; d < DIM; d++) {
            vl2 += L.vel[d] * L.vel[d];
            vr2 += R.vel[d] * R.vel[d];
        }
        const T HL = L.e + vl2 / T(2) + L.p * L.irho;
        const T HR = R.e + vr2 / T(2) + R.p * R.irho;
        flux[kRhoE] = T(0.5) * (HL + HR) * mA;
        for (int d = 0; d < DIM; d++) flux[kRhoU + d] = T(0.5) * (L.vel[d] + R.vel[d]) * mA + p12A * nhat[d];
    }
    mdotA = mA;
    direction = dir;
    rhoL_out = L.rho;
    rhoR_out = R.rho;
}

// Species advection coefficients F_sp = cl rhoY_L + cr rhoY_R from the mass flux mA (area included):
// SpeciesTransport::AdvectionFlux takes the left state for LEFT and the right one otherwise
// (speciesTransport.cpp:325-337); CompactCompressibleNSSpeciesTransport::AdvectionFlux averages rhoY and rho for
// NA (compactCompressibleNSSpeciesTransport.cpp:254-265).
template <typename T>
__device__ inline void species_coefficients(int dir, bool compact, T mA, T rl, T rr, T& cl, T& cr) {
    cl = T(0);
    cr = T(0);
    if (dir == kLeft) {
        cl = mA / rl;
    } else if (dir == kRight || !compact) {
        cr = mA / rr;
    } else {
        cl = mA * T(0.5) / (T(0.5) * (rl + rr));
        cr = cl;
    }
}

// Rusanov full-vector flux for the euler components (build-defined, absent from the reference):
// F = A [ 1/2 (f(U_L) + f(U_R)).n - 1/2 lambda (U_R - U_L) ], lambda = max(|u_n| + a).
template <typename T, int DIM>
__device__ inline void rusanov_flux(const T* uL, const T* uR, const T* nhat, T area, const Phys& ph, T* flux, T& unL,
                                    T& unR, T& lambda) {
    SideState<T, DIM> L, R;
    decode_side<T, DIM>(uL, nhat, ph, L);
    decode_side<T, DIM>(uR, nhat, ph, R);
    lambda = pmax(fabs(L.un) + L.a, fabs(R.un) + R.a);
    unL = L.un;
    unR = R.un;
    const T hl = T(0.5) * lambda;
    flux[kRho] = area * (T(0.5) * (uL[kRho] * L.un + uR[kRho] * R.un) - hl * (uR[kRho] - uL[kRho]));
    flux[kRhoE] = area * (T(0.5) * ((uL[kRhoE] + L.p) * L.un + (uR[kRhoE] + R.p) * R.un) - hl * (uR[kRhoE] - uL[kRhoE]));
    for (int d = 0; d < DIM; d++) {
        const int c = kRhoU + d;
        const T fl = uL[c] * L.un + L.p * nhat[d];
        const T fr = uR[c] * R.un + R.p * nhat[d];
        flux[c] = area * (T(0.5) * (fl + fr) - hl * (uR[c] - uL[c]));
    }
}

// Per-cell viscous terms: the symmetric stress tensor (DIM(DIM+1)/2 entries) and the heat flux k grad T (DIM).
// The face-interpolant viscous flux uses the face average 1/2 (G_L + G_R) of unlimited cell gradients
// (stencils/leastSquaresAverage.cpp:5-67); tau and k grad T are linear in G, so the face values are the averages
// of the per-cell tau and q.
template <int DIM>
struct ViscTerms {
    static constexpr int NTAU = DIM * (DIM + 1) / 2;
    static constexpr int NT = NTAU + DIM;
};
__host__ __device__ constexpr int sym_index(int dim, int c, int d) {
    return c <= d ? c * dim - c * (c - 1) / 2 + (d - c) : d * dim - d * (d - 1) / 2 + (c - d);
}

// CompressibleFlowComputeStressTensor (navierStokesTransport.cpp:510-532) and the conduction term of
// DiffusionFlux (:496-497) for one cell gradient: gA[0][d] = dT/dx_d, gA[1 + c][d] = du_c/dx_d.
template <typename T, int DIM>
__device__ inline void stress_heat(T mu, T k, const T (*gA)[DIM], T* tq) {
    T divVel = T(0);
#pragma unroll
    for (int c = 0; c < DIM; c++) divVel += gA[1 + c][c];
#pragma unroll
    for (int c = 0; c < DIM; c++)
#pragma unroll
        for (int d = c; d < DIM; d++)
            tq[sym_index(DIM, c, d)] =
                (c == d) ? T(2) * mu * (gA[1 + c][d] - divVel / T(3)) : mu * (gA[1 + c][d] + gA[1 + d][c]);
#pragma unroll
    for (int d = 0; d < DIM; d++) tq[ViscTerms<DIM>::NTAU + d] = k * gA[0][d];
}

// NavierStokesTransport::DiffusionFlux (navierStokesTransport.cpp:454-508) from face tau/q and face velocity;
// normal = area-weighted normal.  Adds into flux.
template <typename T, int DIM>
__device__ inline void viscous_flux_add(const T* normal, const T* vel, const T* tq, T* flux) {
#pragma unroll
    for (int c = 0; c < DIM; c++) {
        T vf = T(0);
#pragma unroll
        for (int d = 0; d < DIM; d++) vf += -normal[d] * tq[sym_index(DIM, c, d)];
        flux[kRhoU + c] += vf;
    }
    T e = T(0);
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        T hf = T(0);
#pragma unroll
        for (int c = 0; c < DIM; c++) hf += vel[c] * tq[sym_index(DIM, d, c)];
        hf += tq[ViscTerms<DIM>::NTAU + d];
        e += -normal[d] * hf;
    }
    flux[kRhoE] += e;
}

// PETSc symmetric-form scalar limiters (the PetscLimiterLimit_* set used by cellInterpolant.cpp:316-324).
template <typename T>
__device__ inline T limiter_phi(int type, T f) {
    switch (type) {
        case 0: return T(1);
        case 1: return T(0);
        case 2: return T(2) * pmax(T(0), pmin(f, T(1) - f));
        case 3: return pmax(T(0), T(4) * f * (T(1) - f));
        case 4: return pmax(T(0), T(2) * f * (T(1) - f) / (f * f + (T(1) - f) * (T(1) - f)));
        case 5: return T(2) * pmax(T(0), pmax(pmin(T(2) * f, T(1) - f), pmin(f, T(2) * (T(1) - f))));
        case 6: return pmin(T(1), T(4) * pmax(T(0), pmin(f, T(1) - f)));
        case 7: return sin(T(3.14159265358979323846) * pmax(pmin(f, T(1)), T(0)));
        default: return T(1);
    }
}

}  // namespace abl
