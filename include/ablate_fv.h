/*
 * ablate_fv.h -- C ABI of the MI355X (gfx950) finite-volume compressible-flow right-hand side.
 *
 * The library replaces the per-face CPU loops behind
 *   ablate::finiteVolume::FiniteVolumeSolver::ComputeRHSFunction(PetscReal time, Vec locX, Vec locF)
 *   (src/finiteVolume/finiteVolumeSolver.cpp:233-291, declared finiteVolumeSolver.hpp:106)
 * for the process set NavierStokesTransport (+ SpeciesTransport) with a PerfectGas EOS, a Constant or Sutherland
 * transport model and the Ausm / AverageFlux / OffFlux / AusmpUp flux calculators (plus the build-defined Rusanov
 * flux).
 *
 * Conventions (mirroring PETSc / ABLATE):
 *   - Every entry point returns an int status: 0 = success (PetscErrorCode semantics), nonzero = error;
 *     abl_fv_last_error() returns the message of the last failure on that handle (or the global one when the
 *     handle is NULL).  No C++ exception crosses this boundary.
 *   - State and residual arrays are DEVICE pointers in structure-of-arrays layout U[comp * n_cells + cell].
 *     Component order is the reference's field order: euler {rho, rhoE, rhoU, rhoV[, rhoW]} then densityYi
 *     {rhoY_0 .. rhoY_{ns-1}} (src/finiteVolume/compressibleFlowFields.hpp:16, compressibleFlowFields.cpp:11-60).
 *   - Mesh descriptors are HOST pointers (setup is host-side, like DMPlexComputeGeometryFVM).
 *   - Boundary ghost cells (DM_POLYTOPE_FV_GHOST) occupy the trailing cell range [n_interior, n_cells).
 *   - A handle is thread-compatible, not thread-safe: calls on one handle must be serialised by the caller.
 *   - Results are deterministic: no floating-point atomics; every cell gathers its faces in a fixed order.
 */
#ifndef ABLATE_FV_H
#define ABLATE_FV_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct abl_fv_solver* abl_fv_handle;
typedef struct abl_fv_mesh* abl_fv_mesh_handle;
/* hipStream_t without pulling in the HIP headers; pass NULL for the default stream. */
typedef void* abl_fv_stream;

/* ---- enumerations (values are part of the ABI) ------------------------------------------------------------ */

/* Arithmetic form of the RHS.  The mixed form has no perturbation form of the AusmpUp calculator: abl_fv_create
 * rejects a mixed handle with AusmpUp (create the FP64 form instead). */
enum { ABL_FV_FORM_FP64 = 0, ABL_FV_FORM_MIXED = 1 };

/* Flux calculators: fluxCalculator::Ausm (ausm.cpp:2-43), AverageFlux (averageFlux.cpp:2-10),
 * OffFlux (offFlux.cpp:2-8), AusmpUp (ausmpUp.cpp:5-97; mInf from the physics descriptor, the pressure-gradient
 * scaling alpha from abl_fv_set_pgs); RUSANOV is build-defined (absent from the reference). */
enum {
    ABL_FV_FLUX_NONE = 0,
    ABL_FV_FLUX_AUSM = 1,
    ABL_FV_FLUX_AVERAGE = 2,
    ABL_FV_FLUX_OFF = 3,
    ABL_FV_FLUX_RUSANOV = 4,
    ABL_FV_FLUX_AUSMPUP = 5
};

/* Cell reconstruction (PETSc petscfv_type): upwind = first order, leastsquares = LSQ gradients. */
enum { ABL_FV_RECON_NONE = 0, ABL_FV_RECON_LEASTSQUARES = 1 };

/* Gradient limiters (PETSc petsclimiter_type, symmetric form of cellInterpolant.cpp:282-328) plus the
 * build-defined Barth-Jespersen limiter. */
enum {
    ABL_FV_LIMITER_NONE = 0,
    ABL_FV_LIMITER_ZERO = 1,
    ABL_FV_LIMITER_MINMOD = 2,
    ABL_FV_LIMITER_VANLEER = 3,
    ABL_FV_LIMITER_VANALBADA = 4,
    ABL_FV_LIMITER_SUPERBEE = 5,
    ABL_FV_LIMITER_MC = 6,
    ABL_FV_LIMITER_SIN = 7,
    ABL_FV_LIMITER_BARTH_JESPERSEN = 8
};

/* Transport models (eos/transport): Constant{k, mu, diff} (constant.cpp:3-71) or Sutherland (sutherland.cpp:9-122):
 * mu(T) = 1.716e-5 (T / 273)^(3/2) (273 + 111) / (T + 111), k = mu cp / 0.707, D = mu / (rho 0.707), evaluated at
 * each face from the face temperature and density as NavierStokesTransport::DiffusionFlux and
 * SpeciesTransport::DiffusionSpeciesFlux evaluate them (navierStokesTransport.cpp:454-470,
 * speciesTransport.cpp:196-228).  Sutherland runs in the FP64 form, forward only. */
enum { ABL_FV_TRANSPORT_CONSTANT = 0, ABL_FV_TRANSPORT_SUTHERLAND = 1 };
/* Sutherland{enabledProperties}: properties left out return zero (sutherland.cpp:88-91). */
enum { ABL_FV_TRANSPORT_CONDUCTIVITY = 1, ABL_FV_TRANSPORT_VISCOSITY = 2, ABL_FV_TRANSPORT_DIFFUSIVITY = 4 };

/* Boundary (ghost-cell) conditions.  Every slot that has boundary faces needs one: abl_fv_rhs, abl_fv_rhs_vjp and
 * abl_fv_ts_euler fail while such a slot is still ABL_FV_BC_NONE (the reference would leave those ghost cells
 * unset). */
enum {
    ABL_FV_BC_NONE = 0,
    ABL_FV_BC_INFLOW = 1,  /* EssentialGhost (essentialGhost.cpp:5-23) with a constant (rho, u, T, Yi) state */
    ABL_FV_BC_OUTFLOW = 2, /* zero-gradient extrapolation ghost: U_ghost = U_interior (build-defined) */
    ABL_FV_BC_GHOST = 3    /* caller-filled ghost cells: the ghost rows of U are used exactly as the caller wrote them,
                              as the reference's TS callback inserts boundary values into locX before
                              ComputeRHSFunction (timeStepper.cpp:347-349) with any Ghost update function
                              f(time, face centroid, normal, interior) (ghost.cpp:17-28) */
};

/* ---- descriptors ------------------------------------------------------------------------------------------ */

/* DMPlex-style FV geometry (PetscFVFaceGeom / PetscFVCellGeom, DMPlexComputeGeometryFVM).
 * All vectors are stored with 3 entries even when dim == 2 (the third is ignored). */
typedef struct abl_fv_mesh_desc {
    int32_t dim;                 /* 2 or 3 */
    int64_t n_cells;             /* interior + ghost cells */
    int64_t n_interior;          /* cells [0, n_interior) are interior; [n_interior, n_cells) are boundary ghosts */
    int64_t n_faces;             /* faces with two supporting cells (interior and boundary faces) */
    const int32_t* face_cells;   /* [n_faces][2] support: face_cells[2f] = left (support[0]), [2f+1] = right */
    const double* face_normal;   /* [n_faces][3] area-weighted normal pointing left -> right */
    const double* face_centroid; /* [n_faces][3] face centroid in the frame of the left cell */
    const double* face_shift;    /* [n_faces][3] translation mapping the right cell into the left cell's frame
                                    (periodic images); NULL = all zero */
    const double* cell_centroid; /* [n_cells][3] */
    const double* cell_volume;   /* [n_cells] */
    const int32_t* face_bc;      /* [n_faces] boundary-condition slot of a boundary face (right cell is a ghost),
                                    -1 for interior faces; NULL = no boundary conditions */
} abl_fv_mesh_desc;

/* Physics of one NavierStokesTransport (+ SpeciesTransport) process set.  Key names follow the YAML surface:
 * PerfectGas{gamma, Rgas} (perfectGas.cpp:4-7), Constant{k, mu, diff} (constant.cpp:3-5),
 * NavierStokesTransport{cfl} (navierStokesTransport.cpp:20), CompressibleFlowSolver{maxLimGrad}
 * (compressibleFlowSolver.cpp:59-68), petscfv_type / petsclimiter_type. */
typedef struct abl_fv_physics_desc {
    double gamma;            /* default 1.4 */
    double Rgas;             /* default 287 */
    int32_t n_species;       /* transported species (densityYi components) */
    int32_t flux_calculator; /* ABL_FV_FLUX_* ; NONE = no advection */
    int32_t reconstruction;  /* ABL_FV_RECON_* (applies to euler and densityYi) */
    int32_t limiter;         /* ABL_FV_LIMITER_* */
    double max_lim_grad;     /* maxLimGrad (finiteVolumeSolver.hpp:84 default Constants::large = 1e10) */
    double k;                /* thermal conductivity */
    double mu;               /* dynamic viscosity */
    double diff;             /* species diffusivity */
    double cfl;              /* advection CFL (default 0.5) */
    double ausmpup_minf;     /* AusmpUp{mInf} reference Mach number (ausmpUp.cpp:3; 0 = none) */
    int32_t compact_species; /* 1: CompactCompressibleNSSpeciesTransport (compactCompressibleNSSpeciesTransport.cpp)
                                instead of NavierStokesTransport + SpeciesTransport: NA flux-calculator directions
                                advect the averaged species, the CFL step divides a by the PGS alpha; with PerfectGas
                                (species enthalpies 0) the species-enthalpy diffusion term is zero */
    int32_t transport;          /* ABL_FV_TRANSPORT_CONSTANT (k, mu, diff above) or ABL_FV_TRANSPORT_SUTHERLAND (k,
                                   mu, diff ignored) */
    int32_t transport_disabled; /* Sutherland: ABL_FV_TRANSPORT_* bits of the properties left out of
                                   enabledProperties; 0 = all enabled (the reference's default) */
    double conduction_stability_factor; /* NavierStokesTransport{parameters{conductionStabilityFactor}}: > 0
                                           registers the conduction time step (navierStokesTransport.cpp:31,67-75) */
    double viscous_stability_factor;    /* {viscousStabilityFactor}: > 0 registers the viscous time step (:32,76-78) */
} abl_fv_physics_desc;

/* One boundary-condition slot. */
typedef struct abl_fv_bc_desc {
    int32_t kind;            /* ABL_FV_BC_* */
    double rho;              /* inflow density */
    double vel[3];           /* inflow velocity */
    double T;                /* inflow temperature */
    const double* Yi;        /* inflow mass fractions: exactly n_species values (host pointer, copied), or NULL =
                                zeros */
    int32_t enforce_at_face; /* EssentialGhost{enforceAtFace} (essentialGhost.cpp:15-21): the ghost is
                                2 * state - interior, so that the face average equals the inflow state */
} abl_fv_bc_desc;

/* Read-only facts about a created solver. */
typedef struct abl_fv_info {
    int32_t dim;
    int32_t n_comp;              /* dim + 2 + n_species */
    int32_t max_faces_per_cell;  /* width of the cell -> face table */
    int32_t form;
    int64_t n_cells, n_interior, n_faces;
    int64_t device_bytes;        /* device memory held by the handle (mesh + workspace) */
    double alg_bytes_per_cell;   /* algorithmic HBM bytes per RHS per interior cell (state + mesh) */
} abl_fv_info;

/* ---- mesh generation / ordering (domain helpers; replace BoxMesh + GhostBoundaryCells, boxMesh.cpp:17-43,
 *      modifiers/ghostBoundaryCells.cpp:4-8, and DMPlex RCM reordering) --------------------------------------- */

/* Ordering of generated cells. */
enum {
    ABL_FV_ORDER_NATURAL = 0,
    ABL_FV_ORDER_BRICK = 1,  /* boxes: 8^3 / 16^2 bricks in Morton order of the bricks */
    ABL_FV_ORDER_RCM = 2,    /* reverse Cuthill-McKee of the cell graph */
    ABL_FV_ORDER_MORTON = 3, /* Morton (Z-order) curve of the cell centroids (tet and Gmsh meshes) */
    ABL_FV_ORDER_COLUMN = 4  /* 3-D boxes: columns of 8 x 8 cells in x, y (x-major order of the columns), each column
                                walked in z one 8 x 8 layer after the other (the MARCH strategy's layout) */
};

/* Uniform quad (dim 2) / hex (dim 3) box: cells[d] cells per direction on [lower, upper], periodic[d] != 0 for
 * periodic directions.  Non-periodic directions get one ghost cell per boundary face; boundary faces carry
 * face_bc = 2*d (lower side) or 2*d+1 (upper side).  Faces are numbered by their left cell. */
int abl_fv_mesh_box(int32_t dim, const int32_t* cells, const double* lower, const double* upper, const int32_t* periodic,
                    int32_t ordering, abl_fv_mesh_handle* out);
/* Unit-cube tetrahedral mesh: n^3 cubes split into 6 tets each, interior nodes jittered by up to
 * jitter * h (uniform, seeded), boundary faces with ghost cells (face_bc = 0..5 by cube side).  ordering: natural,
 * RCM or Morton. */
int abl_fv_mesh_tet_box(int32_t n, double jitter, uint64_t seed, int32_t ordering, abl_fv_mesh_handle* out);
/* Mesh file ingestion for `!ablate::domain::MeshFile` inputs in Gmsh format (PETSc's DMPlexCreateGmshFromFile
 * + GhostBoundaryCells + DMPlexComputeGeometryFVM, src/domain/meshFile.cpp): MSH 4.1 ASCII, first-order
 * triangles / quadrilaterals (2-D) or tetrahedra / hexahedra / prisms / pyramids (3-D); boundary faces carry the
 * physical tag of the element on them as face_bc (the "Face Sets" label value; tags must be below 16) and one
 * ghost cell each.  ordering: ABL_FV_ORDER_NATURAL (file order), ABL_FV_ORDER_RCM or ABL_FV_ORDER_MORTON. */
int abl_fv_mesh_gmsh(const char* path, int32_t ordering, abl_fv_mesh_handle* out);
/* Mesh from a DMPlex topology (an interpolated plex, points 0..n_points-1), as an ABLATE domain holds it after
 * DMPlexInterpolate (DMPlexGetCone / DMGetCoordinatesLocal / DMGetLabel("Face Sets"), src/domain/domain.cpp) and as
 * DMPlexView_HDF5 stores it in the Hdf5Serializer output (src/io/hdf5Serializer.cpp:23-160):
 *   cone_offsets[n_points+1], cones   the cone of point p is cones[cone_offsets[p] .. cone_offsets[p+1]);
 *   vertex_coords [nv][dim]           coordinates of the depth-0 points in increasing point number;
 *   face_sets [n_points]              "Face Sets" value of each point (-1: none; NULL when no face is on the boundary).
 * Cells are the points of depth dim (ghost cells of DMPlexConstructGhostCells, with a one-face cone, are dropped and
 * rebuilt); boundary faces take their "Face Sets" value as face_bc (below 16) and one ghost cell each; geometry as
 * DMPlexComputeGeometryFVM gives it for straight-sided convex cells with planar faces (src/solver/cellSolver.cpp:
 * 193-196).  ordering: natural (point order), RCM or Morton. */
int abl_fv_mesh_dmplex(int32_t dim, int64_t n_points, const int64_t* cone_offsets, const int64_t* cones,
                       const double* vertex_coords, const int32_t* face_sets, int32_t ordering,
                       abl_fv_mesh_handle* out);
/* Producer's index of every interior cell (n_interior entries, in the mesh's cell order): the DMPlex point number
 * (abl_fv_mesh_dmplex), the cell element's ordinal in the file (abl_fv_mesh_gmsh), the lexicographic index with x
 * fastest (abl_fv_mesh_box) or the generation index (abl_fv_mesh_tet_box).  Maps fields stored in the producer's
 * cell order onto the mesh. */
int abl_fv_mesh_source_cells(abl_fv_mesh_handle mesh, int64_t* source);
/* Descriptor view of a generated mesh (pointers stay valid until abl_fv_mesh_free). */
int abl_fv_mesh_describe(abl_fv_mesh_handle mesh, abl_fv_mesh_desc* desc);
void abl_fv_mesh_free(abl_fv_mesh_handle mesh);

/* ---- solver lifecycle -------------------------------------------------------------------------------------- */

/* Builds the device mesh (cell->face table, LSQ weights, geometry in the form's precision) and workspace.
 * Replaces FiniteVolumeSolver::Setup/Initialize + the lazy CellInterpolant/FaceInterpolant construction
 * (finiteVolumeSolver.cpp:28-231, cellInterpolant.cpp:5-33,641-889, faceInterpolant.cpp:7-73). */
int abl_fv_create(const abl_fv_mesh_desc* mesh, const abl_fv_physics_desc* physics, int32_t form, abl_fv_handle* out);
int abl_fv_destroy(abl_fv_handle h);
const char* abl_fv_last_error(abl_fv_handle h);
int abl_fv_get_info(abl_fv_handle h, abl_fv_info* info);
/* Registers the state of boundary-condition slot `slot` (Ghost/EssentialGhost, ghost.cpp:17-28). */
int abl_fv_set_boundary(abl_fv_handle h, int32_t slot, const abl_fv_bc_desc* bc);

/* ---- the hot path ------------------------------------------------------------------------------------------- */

/* One RHS evaluation: ghost fill of the INFLOW / OUTFLOW slots (writing those ghost rows of U, the reference's
 * DMPlexInsertBoundaryValues step; ghost rows of ABL_FV_BC_GHOST slots are read as the caller filled them and
 * never written), aux decode, gradients + limiter, face fluxes and per-cell residual gather.  With accumulate = 1
 * the residual is ADDED into R (reference semantics: locF is zeroed by the caller, timeStepper.cpp:363, and
 * ComputeRHSFunction adds into it); with accumulate = 0, R = rhs.  Ghost rows of R are left untouched
 * (accumulate = 1) or zeroed.  `time` is the time the caller's ghost values belong to; the constant-state
 * conditions do not depend on it.  Fails before any device work when a boundary slot has no condition. */
int abl_fv_rhs(abl_fv_handle h, double time, double* U, double* R, int32_t accumulate, abl_fv_stream stream);

/* Vector-Jacobian product of the RHS (build-defined; the reference has no adjoint):
 *   Ubar = Rbar^T dR/dU   (interior rows; the ghost rows of ABL_FV_BC_GHOST slots receive the adjoint of the
 *   caller's ghost values, which are inputs of the RHS; other ghost rows receive zero, their dependence on the
 *   interior (OUTFLOW, enforce_at_face) and on theta being folded into Ubar and theta_bar)
 *   theta_bar[3 * slot + {0,1,2}] += Rbar^T dR/d(rho_in, u_in, T_in) for every INFLOW slot (host array sized
 *   3 * number of slots, or NULL; with inflow slots the call waits for the stream to copy their sums back, without
 *   any it leaves theta_bar as it is and returns without waiting).  Fluxes are recomputed, not stored.  Supported
 *   for the FP64 form with the
 *   Ausm, Rusanov or Off calculators, first order or least squares with the Barth-Jespersen limiter (or none),
 *   viscous and species fluxes; the first call allocates its workspace (first order: 2 n_euler + 2 + n_species
 *   doubles per face). */
int abl_fv_rhs_vjp(abl_fv_handle h, double time, double* U, const double* Rbar, double* Ubar, double* theta_bar,
                   abl_fv_stream stream);

/* FaceInterpolant::GetInterpolatedFaceVectors (faceInterpolant.cpp:109-249, the public method the reference's
 * faceInterpolantTests.cpp:224-265 checks): per face f, the face values and gradients the continuous (viscous,
 * conduction, species-diffusion) fluxes use, from the least-squares-average stencils
 * (stencils/leastSquaresAverage.cpp:5-155): value 1/2 (q_L + q_R), gradient 1/2 (G_L + G_R) of unlimited
 * least-squares cell gradients over all face neighbours (a ghost cell's: w (q_interior - q_ghost)).  Solution
 * components q < n_comp into face_sol [n_comp][n_faces] and face_sol_grad [n_comp * dim + d][n_faces]; aux fields
 * (temperature, velocity, Yi; 1 + dim + n_species) into face_aux / face_aux_grad likewise.  Device pointers; any
 * output may be NULL.  Ghost rows of U are used as they stand.  FP64 handles with a transport term only. */
int abl_fv_face_interpolate(abl_fv_handle h, const double* U, double* face_sol, double* face_sol_grad,
                            double* face_aux, double* face_aux_grad, abl_fv_stream stream);

/* No counterpart in the reference: abl_fv_rhs_vjp for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  Same arguments and results as
 * abl_fv_rhs_vjp (bitwise), for first-order inviscid FP64 handles with at most 262144 boundary faces; it allocates
 * nothing: the theta_bar workspace lives in a static per-GPU device array.  Calls that overlap on one GPU race on
 * that array: serialise them.  Fails (nonzero, message in abl_fv_last_error) for other handles or larger meshes. */
int abl_fv_rhs_vjp_noalloc(abl_fv_handle h, double time, double* U, const double* Rbar, double* Ubar,
                           double* theta_bar, abl_fv_stream stream);

/* DMPlex local-vector layout (AoS per cell, fields contiguous) <-> SoA. Device pointers. */
int abl_fv_pack_aos(abl_fv_handle h, const double* aos, double* soa, abl_fv_stream stream);
int abl_fv_unpack_aos(abl_fv_handle h, const double* soa, double* aos, abl_fv_stream stream);

/* Number of interior cells whose state holds a non-finite value or a non-positive density
 * (mirrors the verboseSourceCheck scan, timeStepper.cpp:385-389). */
int abl_fv_check_state(abl_fv_handle h, const double* U, int64_t* n_bad, abl_fv_stream stream);

/* ---- pressure-gradient scaling (processes::PressureGradientScaling, pressureGradientScaling.cpp:7-137) ------- */

/* Constructor arguments of PressureGradientScaling{alphaInit, domainLength, maxAlphaAllowed, maxDeltaPressureFac}
 * (pressureGradientScaling.cpp:7-14: maxAlphaAllowed < 1 means 100, maxDeltaPressureFac <= 0 means 0.05). */
typedef struct abl_fv_pgs_desc {
    double alpha_init;
    double domain_length;
    double max_alpha_allowed;
    double max_delta_pressure_fac;
} abl_fv_pgs_desc;
/* Attaches a PressureGradientScaling to the handle (NULL detaches): alpha starts at alpha_init, and the AusmpUp flux
 * calculator uses it (AusmpUp{pgs}, ausmpUp.cpp:9). */
int abl_fv_set_pgs(abl_fv_handle h, const abl_fv_pgs_desc* pgs);
/* PressureGradientScaling::UpdatePreconditioner (pressureGradientScaling.cpp:16-137), the pre-step update: pressure
 * mean / extremes, maximum sound speed and Mach number over the interior cells of U (device reductions), then the
 * reference's alpha update with the step size dt.  Returns the new alpha and the maximum Mach number (host values;
 * synchronises the stream).  Fails when no PressureGradientScaling is attached. */
int abl_fv_pgs_update(abl_fv_handle h, const double* U, double dt, double* alpha, double* max_mach,
                      abl_fv_stream stream);
/* GetAlpha / restore of the saved alpha (pressureGradientScaling.cpp:154-164). */
int abl_fv_get_pgs_alpha(abl_fv_handle h, double* alpha);
int abl_fv_set_pgs_alpha(abl_fv_handle h, double alpha);

/* ---- time step (the reference's CFL step size and forward-Euler stage, on the device) ------------------------- */

/* NavierStokesTransport::ComputeCflTimeStep (navierStokesTransport.cpp:209-285, pgsAlpha = 1; with compact_species
 * CompactCompressibleNSSpeciesTransport::ComputeCflTimeStep, a divided by the attached PGS alpha): the minimum over
 * interior cells of cfl * 2 r_min / (a + sum_d |rhoU_d / rho|), r_min the MIN_CELL_RADIUS mesh characteristic
 * (finiteVolumeSolver.cpp:185-215), cfl from the physics descriptor.  *dt is a host value (synchronises the
 * stream).  Cells with a NaN step are skipped. */
int abl_fv_cfl_dt(abl_fv_handle h, const double* U, double* dt, abl_fv_stream stream);
/* FiniteVolumeSolver::ComputePhysicsTimeSteps (finiteVolumeSolver.cpp:381-393) of the process set: dt[0] the CFL
 * step (registered with a flux calculator, navierStokesTransport.cpp:37-43), dt[1] ComputeConductionTimeStep
 * (:287-371: f_cond (2 r_min)^2 rho cv / k per cell) and dt[2] ComputeViscousDiffusionTimeStep (:373-450:
 * f_visc (2 r_min)^2 rho / mu per cell), each the minimum over interior cells starting from Constants::large = 1e10,
 * with k and mu of the transport model at the cell temperature.  Bit i of *registered is set when function i is
 * registered (the diffusion steps when their stability factor is > 0).  Host values (synchronises the stream). */
int abl_fv_time_steps(abl_fv_handle h, const double* U, double* dt, int32_t* registered, abl_fv_stream stream);
/* FiniteVolumeSolver::ComputePhysicsTimeStep (finiteVolumeSolver.cpp:371-379): the minimum of the registered
 * time-step functions of abl_fv_time_steps (1e10 when none is registered).  Host value (synchronises). */
int abl_fv_physics_dt(abl_fv_handle h, const double* U, double* dt, abl_fv_stream stream);
/* Forward-Euler stage in FP64 on the interior rows: U += dt * R (ghost rows untouched). */
int abl_fv_axpy(abl_fv_handle h, double dt, const double* R, double* U, abl_fv_stream stream);
/* Integrates U from t0 to max_time with forward Euler as the reference's TimeStepper runs it under
 * ts_adapt_type physics (safety factor `safety`), ts_exact_final_time matchstep and PETSc's TSEULER step
 * (timeStepper.cpp:80-130, adaptPhysics.cpp:3-59): first step min(dt0, safety * physics_dt(U0)), the physics step
 * being abl_fv_physics_dt (CFL and, when registered, the diffusion steps); each step evaluates
 * the RHS into the caller's workspace R [n_comp][n_cells], takes the next step size from U_n, updates U; steps
 * that would pass max_time are cut to it.  Stops at max_time or after max_steps; reports the final time and
 * the number of steps taken.  With a PressureGradientScaling attached, its pre-step update (abl_fv_pgs_update with the
 * step about to be taken, registered by PressureGradientScaling::Setup as a TS pre-step, pressureGradientScaling.cpp:
 * 144-146) runs before every step. */
int abl_fv_ts_euler(abl_fv_handle h, double* U, double* R, double t0, double max_time, double dt0, double safety,
                    int64_t max_steps, double* t_out, int64_t* steps_out, abl_fv_stream stream);

/* ---- evaluation strategy (all deterministic, no float atomics).  The reference has one CPU loop order; these
 *      have no counterpart there and all give the residual of ComputeRHSFunction. ------------------------------ */
enum {
    ABL_FV_STRATEGY_AUTO = 0,      /* pick TILE or FACE_FLUX from the mesh's tile statistics (the default) */
    ABL_FV_STRATEGY_RECOMPUTE = 1, /* one lane per cell evaluates each of its faces: every face flux computed twice
                                      (mixed form: no species) */
    ABL_FV_STRATEGY_FACE_FLUX = 2, /* one lane per face stores its flux once; a second pass gathers per cell
                                      (FP64: 8 * (n_euler + n_species rounded up to even) bytes per face; mixed:
                                      8 * (n_euler + n_species)) */
    ABL_FV_STRATEGY_TILE = 3,      /* one workgroup per tile of consecutive cells evaluates the tile's faces once into
                                      LDS and gathers its cells; bitwise equal to FACE_FLUX */
    ABL_FV_STRATEGY_FUSED = 4,     /* one workgroup per tile computes the gradients of its cells and of its halo into
                                      LDS (least-squares normal equations solved per cell, no stored weights), then
                                      the tile's face fluxes and the gather: no cell record or face flux in HBM.
                                      FP64 form without species, Constant transport, no limiter or Barth-Jespersen,
                                      full-rank stencils */
    ABL_FV_STRATEGY_MARCH = 5      /* one workgroup per sequence of 64-cell tiles that share faces (the z-layers of a
                                      column in ABL_FV_ORDER_COLUMN): each step computes the records of the next tile
                                      into LDS, evaluates the current tile's faces and gathers it, so faces between
                                      consecutive tiles are evaluated once; same availability as FUSED */
};
/* AUTO chooses TILE when the tile kernel evaluates at most ~1.8 faces per face
 * (border faces twice) and its LDS footprint leaves two or more waves per SIMD, FACE_FLUX otherwise.  Selecting FACE_FLUX allocates its face-flux
 * storage here, never inside abl_fv_rhs; TILE fails when no tile size fits the LDS budget. */
int abl_fv_set_strategy(abl_fv_handle h, int32_t strategy);
/* Cells per tile of the tile strategy: 0 = the default choice (the largest of 256/128/64 whose face fluxes fit
 * 64 KB of LDS), or exactly 64, 128, 256 or 512 (up to the 160 KB a gfx950 workgroup may hold).  Fails when the
 * tile's face fluxes do not fit.  The residual does not depend on it (bitwise). */
int abl_fv_set_tile_size(abl_fv_handle h, int32_t cells);
/* The strategy abl_fv_rhs runs (AUTO resolved). */
int abl_fv_get_strategy(abl_fv_handle h, int32_t* strategy);

/* ---- per-phase timing (replaces the Loggable / PetscLogEvent brackets around each RHS sub-phase,
 *      finiteVolumeSolver.cpp:239-288, src/utilities/loggable.hpp:26-48) ------------------------------------- */
enum {
    ABL_FV_PHASE_GHOST = 0,
    ABL_FV_PHASE_GRADIENTS = 1,
    ABL_FV_PHASE_SPECIES_GRADIENTS = 2,
    ABL_FV_PHASE_FLUX = 3,
    ABL_FV_PHASE_VJP = 4,
    ABL_FV_PHASE_GATHER = 5,
    ABL_FV_N_PHASES = 6
};
/* When enabled, every kernel of abl_fv_rhs / abl_fv_rhs_vjp is bracketed by events on its stream. */
int abl_fv_set_profiling(abl_fv_handle h, int32_t enable);
/* Accumulated device milliseconds and launch counts per phase since the last reset (waits for the events). */
int abl_fv_get_phase_times(abl_fv_handle h, double* ms, int64_t* launches, int32_t reset);

/* ---- measurement utility (no counterpart in the reference) ---------------------------------------------- */

/* Achievable-bandwidth reference: times `reps` streaming passes over a `bytes`-sized device buffer (kind 0: copy
 * into a second buffer, counted as 2 x bytes moved; kind 1: read only) with 16-byte loads per lane, after two
 * warm-up passes; returns GB/s and the milliseconds per pass.  Allocates and frees its buffers. */
int abl_hbm_probe(int32_t kind, int64_t bytes, int32_t reps, double* gb_per_s, double* ms);

#ifdef __cplusplus
}
#endif
#endif /* ABLATE_FV_H */
