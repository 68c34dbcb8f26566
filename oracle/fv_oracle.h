/*
 * fv_oracle.h -- CPU FP64 restatement of ABLATE's finite-volume compressible RHS.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker: tests, the smoke check and the bench's CPU-baseline leg may
 * call it; the product library never links or calls it.  Each function cites the reference file:line it
 * restates (paths relative to the ABLATE source tree).
 *
 * Data model follows the reference: the local state vector is AoS per cell (DMPlex local section, fields
 * contiguous: euler {rho, rhoE, rhoU..} then densityYi), the residual is accumulated into locF with the
 * reference's per-face `fL -= F/V_L; fR += F/V_R` scatter, faces are visited in increasing face number.
 */
#ifndef FV_ORACLE_H
#define FV_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Flux calculator direction, fluxCalculator.hpp:12 */
enum { ORC_NA = 0, ORC_LEFT = 1, ORC_RIGHT = 2 };
/* Same numeric values as the product ABI enums (checked by the tests). */
enum { ORC_FLUX_NONE = 0, ORC_FLUX_AUSM = 1, ORC_FLUX_AVERAGE = 2, ORC_FLUX_OFF = 3, ORC_FLUX_RUSANOV = 4,
       ORC_FLUX_AUSMPUP = 5 };
enum { ORC_RECON_NONE = 0, ORC_RECON_LSQ = 1 };
enum {
    ORC_LIM_NONE = 0, ORC_LIM_ZERO, ORC_LIM_MINMOD, ORC_LIM_VANLEER, ORC_LIM_VANALBADA,
    ORC_LIM_SUPERBEE, ORC_LIM_MC, ORC_LIM_SIN, ORC_LIM_BJ
};
/* ORC_BC_GHOST: ghost values supplied by the caller in locX (a Ghost update function evaluated before the RHS,
 * ghost.cpp:17-28, timeStepper.cpp:347-349): left untouched. */
enum { ORC_BC_NONE = 0, ORC_BC_INFLOW = 1, ORC_BC_OUTFLOW = 2, ORC_BC_GHOST = 3 };
enum { ORC_TRANSPORT_CONSTANT = 0, ORC_TRANSPORT_SUTHERLAND = 1 };
enum { ORC_TRANSPORT_CONDUCTIVITY = 1, ORC_TRANSPORT_VISCOSITY = 2, ORC_TRANSPORT_DIFFUSIVITY = 4 };

typedef struct {
    int dim;
    int64_t n_cells, n_interior, n_faces;
    const int32_t* face_cells;   /* [F][2] */
    const double* face_normal;   /* [F][3] */
    const double* face_centroid; /* [F][3] (left cell frame) */
    const double* face_shift;    /* [F][3] or NULL */
    const double* cell_centroid; /* [N][3] */
    const double* cell_volume;   /* [N] */
    const int32_t* face_bc;      /* [F] or NULL */
} orc_mesh;

typedef struct {
    double gamma, Rgas;
    int n_species;
    int flux;
    int recon;
    int limiter;
    double max_lim_grad;
    double k, mu, diff;
    double ausmpup_minf; /* AusmpUp mInf (0 = none) */
    double pgs_alpha;    /* PressureGradientScaling alpha seen by AusmpUp (<= 0: no PGS, alpha 1) */
    int compact;         /* CompactCompressibleNSSpeciesTransport instead of NavierStokes + SpeciesTransport */
    int transport;       /* ORC_TRANSPORT_CONSTANT (k, mu, diff) or ORC_TRANSPORT_SUTHERLAND */
    int transport_disabled; /* Sutherland: ORC_TRANSPORT_* bits of the properties left out (0 = all enabled) */
} orc_physics;

typedef struct {
    int kind;
    double rho, vel[3], T;
    const double* Yi;
    int enforce_at_face; /* EssentialGhost{enforceAtFace}: ghost = 2 state - interior (essentialGhost.cpp:15-21) */
} orc_bc;

/* ---- leaf functions (pinned by the reference's unit-test tables) ---- */
int orc_ausm(double uL, double aL, double rhoL, double pL, double uR, double aR, double rhoR, double pR, double* massFlux,
             double* p12);
int orc_average_flux(double uL, double aL, double rhoL, double pL, double uR, double aR, double rhoR, double pR,
                     double* massFlux, double* p12);
int orc_off_flux(double uL, double aL, double rhoL, double pL, double uR, double aR, double rhoR, double pR,
                 double* massFlux, double* p12);
/* AusmpUp (ausmpUp.cpp:5-97); pgs_alpha = 1 without a PressureGradientScaling. */
int orc_ausmpup(double mInf, double pgs_alpha, double uL, double aL, double rhoL, double pL, double uR, double aR,
                double rhoR, double pR, double* massFlux, double* p12);

/* PerfectGas property functions on an euler field of `dim` momentum components at offset 0 (perfectGas.cpp). */
double orc_pg_pressure(const double* euler, int dim, double gamma, double R);
double orc_pg_temperature(const double* euler, int dim, double gamma, double R);
double orc_pg_internal_energy(const double* euler, int dim);
double orc_pg_internal_energy_T(double T, double gamma, double R);
double orc_pg_speed_of_sound(const double* euler, int dim, double gamma, double R);
double orc_pg_speed_of_sound_T(const double* euler, double T, double gamma, double R);
double orc_pg_pressure_T(const double* euler, double T, double R);
double orc_pg_sensible_enthalpy_T(double T, double gamma, double R);
/* transport::Sutherland at temperature T, density rho and cp: out = {conductivity, viscosity, diffusivity}
 * (sutherland.cpp:9-64; a property left out of enabledProperties is zero, :88-91). */
void orc_sutherland(double T, double rho, double cp, int disabled, double* out);
/* ComputeConductionTimeStep / ComputeViscousDiffusionTimeStep (navierStokesTransport.cpp:287-450) over the interior
 * cells with stability factors fcond, fvisc and per-cell MIN_CELL_RADIUS r: out = {conduction dt, viscous dt}. */
void orc_diffusion_dt(const orc_mesh* m, const orc_physics* ph, double fcond, double fvisc, const double* r,
                      const double* X, double* out);
double orc_pg_cv(double gamma, double R);
double orc_pg_cp(double gamma, double R);

/* NavierStokesTransport::AdvectionFlux for a face with area normal `normal` (navierStokesTransport.cpp:99-207).
 * fieldL/R = euler components. Writes dim+2 values. */
int orc_ns_advection_flux(int dim, const double* normal, const double* fieldL, const double* fieldR, double gamma,
                          double R, int flux_calculator, double* flux);
/* CompressibleFlowComputeStressTensor (navierStokesTransport.cpp:510-532). */
void orc_stress_tensor(int dim, double mu, const double* gradVel, double* tau);
/* Least-squares pseudo-inverse weights of PetscFVComputeGradient: grad[f*dim+d] = pinv(dx)[d][f]. */
int orc_lsq_weights(int n, int dim, const double* dx, double* grad);
/* Symmetric-form scalar limiter phi(f) (PETSc PetscLimiterLimit_*). */
double orc_limiter(int type, double f);
/* Build-defined Rusanov flux of one face (area-weighted normal), euler then ns species components. */
void orc_rusanov_flux(int dim, const double* normal, const double* uL, const double* uR, double gamma, double R, int ns,
                      double* flux);

/* Full RHS: ghost fill of locX (boundary faces with a bc slot), aux update, discontinuous (cell interpolant) and
 * continuous (face interpolant) flux functions, accumulated into locF (AoS [n_cells][dim+2+ns]).
 * bcs may be NULL when the mesh has no boundary faces. Returns 0 on success. */
int orc_rhs(const orc_mesh* mesh, const orc_physics* phys, const orc_bc* bcs, int n_bcs, double time, double* locX,
            double* locF);

/* Setup + repeated evaluation (the setup holds the LSQ weights and face stencils the reference builds once). */
typedef struct orc_ctx orc_ctx;
int orc_create(const orc_mesh* mesh, const orc_physics* phys, orc_ctx** out);
int orc_rhs_ctx(orc_ctx* ctx, const orc_bc* bcs, int n_bcs, double time, double* locX, double* locF);
void orc_destroy(orc_ctx* ctx);

/* Cell gradients (after limiting) of the euler+densityYi field as the oracle computes them, AoS
 * [n_cells][totDim][dim] (zero for ghost cells). For tests of the reconstruction stage. */
int orc_cell_gradients(const orc_mesh* mesh, const orc_physics* phys, const double* locX, double* grad);

/* FaceInterpolant::GetInterpolatedFaceVectors (faceInterpolant.cpp:109-249) with the LeastSquaresAverage stencils
 * (stencils/leastSquaresAverage.cpp:5-155): per face, the interpolated solution values [F][nc], solution gradients
 * [F][nc][dim], aux values (T, velocity, Yi) [F][naux] and aux gradients [F][naux][dim], naux = 1 + dim + ns.
 * Ghost values are taken from locX as given (no boundary fill). */
int orc_face_interpolate(const orc_mesh* mesh, const orc_physics* phys, const double* locX, double* faceSol,
                         double* faceSolGrad, double* faceAux, double* faceAuxGrad);

/* Minimum cell radius per cell: min over the cell's faces of |x_cell - x_face| (the MIN_CELL_RADIUS mesh
 * characteristic, finiteVolumeSolver.cpp:185-215; every face with two supporting cells counts).  r[n_cells]. */
void orc_min_cell_radius(const orc_mesh* mesh, double* r);

/* NavierStokesTransport::ComputeCflTimeStep (navierStokesTransport.cpp:209-285) with pgsAlpha = 1, or with
 * phys->compact CompactCompressibleNSSpeciesTransport::ComputeCflTimeStep (compactCompressibleNSSpeciesTransport.cpp:
 * 273-345, a divided by the PGS alpha):
 * min over interior cells of cfl * 2 r_min / (a + sum_d |rhoU_d / rho|), a from the PerfectGas T route.
 * locX is AoS [n_cells][dim+2+ns]. */
double orc_cfl_dt(const orc_mesh* mesh, const orc_physics* phys, double cfl, const double* r, const double* locX);

/* PressureGradientScaling::UpdatePreconditioner (pressureGradientScaling.cpp:16-137) over the interior cells of the
 * AoS state locX: updates *alpha with step dt, writes the maximum Mach number. */
void orc_pgs_update(const orc_mesh* mesh, const orc_physics* phys, const double* locX, double dt, double domain_length,
                    double max_alpha_allowed, double max_delta_pressure_fac, double* alpha, double* max_mach);

#ifdef __cplusplus
}
#endif
#endif
