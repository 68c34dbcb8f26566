"""Host-side mirror of ABLATE's finite-volume process / flux-calculator / EOS surface, backed by the HIP RHS.

Class and parameter names follow the reference's YAML surface:
  ablate::eos::PerfectGas{parameters{gamma, Rgas}, species}                 (src/eos/perfectGas.cpp:4-7,328-329)
  ablate::eos::transport::Constant{k, mu, diff}                             (src/eos/transport/constant.cpp:74-75)
  ablate::eos::transport::Sutherland{eos, enabledProperties}                (src/eos/transport/sutherland.cpp:120-122)
  ablate::finiteVolume::fluxCalculator::{Ausm, AverageFlux, OffFlux, AusmpUp{mInf, pgs}}   (fluxCalculator/*.cpp)
  ablate::finiteVolume::processes::PressureGradientScaling{eos, alphaInit, domainLength, maxAlphaAllowed,
                                                           maxDeltaPressureFac}  (processes/pressureGradientScaling.cpp)
  ablate::finiteVolume::processes::NavierStokesTransport{parameters{cfl, conductionStabilityFactor,
    viscousStabilityFactor}, eos, fluxCalculator, transport}
                                                                            (processes/navierStokesTransport.cpp:569-574)
  ablate::finiteVolume::processes::SpeciesTransport{eos, fluxCalculator, transport}   (speciesTransport.cpp:474-478)
  ablate::finiteVolume::FiniteVolumeSolver::ComputeRHSFunction(time, locX, locF)     (finiteVolumeSolver.cpp:233-291)
Rusanov is build-defined (the reference has no Rusanov flux calculator).

State and residual are CUDA (HIP) float64 tensors in structure-of-arrays layout ``[n_comp, n_cells]``.
"""
from __future__ import annotations

import ctypes
import weakref
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .mesh import _NativeMesh


# ----------------------------------------------------------------------------------------------- EOS / transport
@dataclass
class PerfectGas:
    parameters: dict = field(default_factory=dict)
    species: list = field(default_factory=list)

    @property
    def gamma(self) -> float:
        return float(self.parameters.get("gamma", 1.4))

    @property
    def Rgas(self) -> float:
        return float(self.parameters.get("Rgas", 287.0))


@dataclass
class Constant:
    k: float = 0.0
    mu: float = 0.0
    diff: float = 0.0


@dataclass
class Sutherland:
    """transport::Sutherland (sutherland.cpp:9-122): mu(T) = 1.716e-5 (T / 273)^(3/2) 384 / (T + 111),
    k = mu cp / 0.707, D = mu / (rho 0.707), evaluated at each face from its temperature and density.
    ``enabledProperties``: any of "Conductivity", "Viscosity", "Diffusivity"; empty = all (the others are zero).
    FP64 form, forward only."""

    eos: PerfectGas | None = None
    enabledProperties: list = field(default_factory=list)

    BITS = {"conductivity": 1, "viscosity": 2, "diffusivity": 4}

    def disabled_bits(self) -> int:
        if not self.enabledProperties:
            return 0
        names = {str(x).lower() for x in self.enabledProperties}
        unknown = names - set(self.BITS)
        if unknown:
            raise ValueError(f"unknown transport properties {sorted(unknown)}")
        return sum(bit for name, bit in self.BITS.items() if name not in names)


# ----------------------------------------------------------------------------------------------- flux calculators
class FluxCalculator:
    kind = "none"


class Ausm(FluxCalculator):
    """AUSM flux splitting, Liou and Steffen 1993 (fluxCalculator/ausm.cpp:2-43)."""

    kind = "ausm"


class AverageFlux(FluxCalculator):
    kind = "average"


class OffFlux(FluxCalculator):
    kind = "off"


class AusmpUp(FluxCalculator):
    """AUSM+-up, Liou 2006 (fluxCalculator/ausmpUp.cpp:5-97): reference Mach number mInf (0 = none) and an optional
    PressureGradientScaling whose alpha scales the acoustic speed."""

    kind = "ausmpup"

    def __init__(self, mInf: float = 0.0, pgs: "PressureGradientScaling | None" = None):
        self.mInf = float(mInf)
        self.pgs = pgs


class Rusanov(FluxCalculator):
    """Local Lax-Friedrichs flux (build-defined, full-vector)."""

    kind = "rusanov"


# ----------------------------------------------------------------------------------------------- processes
@dataclass
class NavierStokesTransport:
    parameters: dict = field(default_factory=dict)
    eos: PerfectGas = field(default_factory=PerfectGas)
    fluxCalculator: FluxCalculator | None = None
    transport: Constant | Sutherland | None = None
    pgs: object = None


class PressureGradientScaling:
    """processes::PressureGradientScaling (pressureGradientScaling.cpp:7-137): alpha rescales the acoustic speed
    seen by AusmpUp; UpdatePreconditioner(dt) is the pre-step update (device reductions over the interior cells).
    Attached to the solver it is listed in; GetAlpha / GetMaxMach read its state."""

    def __init__(self, eos=None, alphaInit: float = 1.0, domainLength: float = float("nan"),
                 maxAlphaAllowed: float | None = None, maxDeltaPressureFac: float | None = None):
        self.eos = eos
        self.alphaInit = float(alphaInit)
        self.domainLength = float(domainLength)
        self.maxAlphaAllowed = 0.0 if maxAlphaAllowed is None else float(maxAlphaAllowed)
        self.maxDeltaPressureFac = 0.0 if maxDeltaPressureFac is None else float(maxDeltaPressureFac)
        self._solver = lambda: None  # weak reference to the solver it is attached to
        self.maxMach = 0.0

    def desc(self) -> _lib.PgsDesc:
        return _lib.PgsDesc(alpha_init=self.alphaInit, domain_length=self.domainLength,
                            max_alpha_allowed=self.maxAlphaAllowed, max_delta_pressure_fac=self.maxDeltaPressureFac)

    def GetAlpha(self) -> float:
        sv = self._solver()
        if sv is None:
            return self.alphaInit
        a = ctypes.c_double()
        _lib.check(_lib.lib().abl_fv_get_pgs_alpha(sv.handle, ctypes.byref(a)), sv.handle, "abl_fv_get_pgs_alpha")
        return float(a.value)

    def GetMaxMach(self) -> float:
        return self.maxMach

    def UpdatePreconditioner(self, dt: float, U, stream=None) -> float:
        """The pre-step alpha update from the state U (device, SoA); returns the new alpha."""
        sv = self._solver()
        if sv is None:
            raise RuntimeError("PressureGradientScaling is not attached to a FiniteVolumeSolver")
        sv._check_tensor(U, "U")
        a, m = ctypes.c_double(), ctypes.c_double()
        _lib.check(_lib.lib().abl_fv_pgs_update(sv.handle, ctypes.c_void_p(U.data_ptr()), float(dt), ctypes.byref(a),
                                                ctypes.byref(m), sv._stream_ptr(stream)), sv.handle,
                   "abl_fv_pgs_update")
        self.maxMach = float(m.value)
        return float(a.value)


@dataclass
class CompactCompressibleNSSpeciesTransport:
    """processes::CompactCompressibleNSSpeciesTransport (compactCompressibleNSSpeciesTransport.cpp): euler and
    densityYi in one process.  Same fluxes as NavierStokesTransport + SpeciesTransport except that a flux
    calculator's NA direction advects the averaged species and the CFL step divides the sound speed by the `pgs`
    alpha; the species-enthalpy diffusion term is zero for PerfectGas (species sensible enthalpies 0)."""

    parameters: dict = field(default_factory=dict)
    eos: PerfectGas = field(default_factory=PerfectGas)
    fluxCalculator: FluxCalculator | None = None
    transport: Constant | Sutherland | None = None
    pgs: object = None


@dataclass
class SpeciesTransport:
    eos: PerfectGas = field(default_factory=PerfectGas)
    fluxCalculator: FluxCalculator | None = None
    transport: Constant | Sutherland | None = None
    parameters: dict = field(default_factory=dict)


# ----------------------------------------------------------------------------------------------- boundaries
@dataclass
class InflowGhost:
    """EssentialGhost with a constant (rho, velocity, T, Yi) state (boundaryConditions/essentialGhost.cpp);
    ``enforceAtFace`` sets the ghost to 2 state - interior so that the face average takes the state
    (essentialGhost.cpp:15-21)."""

    labelIds: list
    rho: float
    velocity: list
    T: float
    Yi: list | None = None
    enforceAtFace: bool = False


class Ghost:
    """boundaryConditions::Ghost (ghost.cpp:17-28): the ghost value of field `fieldName` on the faces of `labelIds`
    is ``updateFunction(time, centroid, normal, xI) -> xG`` of the face centroid, the area-weighted normal (left to
    right, i.e. out of the domain) and the interior cell's field values, all [n_faces, ...] numpy arrays.  The solver
    evaluates it on the host before each RHS and writes the ghost rows (the reference's boundary-value insertion,
    timeStepper.cpp:347-349); the device path treats those rows as caller-filled (ABL_FV_BC_GHOST)."""

    def __init__(self, fieldName: str, boundaryName: str, labelIds, updateFunction=None):
        """updateFunction None: the caller writes the ghost rows of U itself before each RHS (the C contract)."""
        if fieldName not in ("euler", "densityYi"):
            raise ValueError("Ghost applies to the euler or densityYi field")
        self.fieldName, self.boundaryName = fieldName, boundaryName
        self.labelIds = list(labelIds)
        self.updateFunction = updateFunction


@dataclass
class OutflowGhost:
    """Zero-gradient extrapolation ghost (build-defined)."""

    labelIds: list


# ----------------------------------------------------------------------------------------------- solver
class FiniteVolumeSolver:
    """Drop-in for FiniteVolumeSolver::ComputeRHSFunction on one GPU.

    options: ``petscfv_type`` ("upwind" | "leastsquares") and ``petsclimiter_type`` ("none", "minmod",
    "vanleer", ..., "barthjespersen"); ``maxLimGrad`` as CompressibleFlowSolver{maxLimGrad}.
    """

    def __init__(self, mesh: _NativeMesh, processes, options: dict | None = None, boundaryConditions=(),
                 maxLimGrad: float = 1e10, form: str = "fp64"):
        options = options or {}
        ns_proc = [p for p in processes if isinstance(p, (NavierStokesTransport, CompactCompressibleNSSpeciesTransport))]
        sp_proc = [p for p in processes if isinstance(p, SpeciesTransport)]
        if len(ns_proc) != 1:
            raise ValueError("exactly one NavierStokesTransport or CompactCompressibleNSSpeciesTransport process is "
                             "required")
        nst = ns_proc[0]
        compact = isinstance(nst, CompactCompressibleNSSpeciesTransport)
        if compact and sp_proc:
            raise ValueError("CompactCompressibleNSSpeciesTransport transports the species itself")
        eos = nst.eos
        n_species = len(eos.species)
        flux = nst.fluxCalculator.kind if nst.fluxCalculator else "none"
        transport = nst.transport
        if sp_proc:
            spt = sp_proc[0]
            if (spt.fluxCalculator.kind if spt.fluxCalculator else "none") != flux:
                raise ValueError("SpeciesTransport and NavierStokesTransport must use the same flux calculator")
            if (spt.transport or Constant()) != (transport or Constant()):
                raise ValueError("SpeciesTransport and NavierStokesTransport must use the same transport model")
        elif n_species and not compact:
            raise ValueError("species are declared in the EOS but no SpeciesTransport process was given")
        tr = transport or Constant()
        suth = isinstance(tr, Sutherland)
        if suth and form == "mixed":
            raise ValueError("the mixed form has no Sutherland transport: use form='fp64'")
        recon = options.get("petscfv_type", "upwind")
        limiter = options.get("petsclimiter_type", "none").replace("-", "").replace("_", "")
        phys = _lib.PhysicsDesc(
            gamma=eos.gamma,
            Rgas=eos.Rgas,
            n_species=n_species,
            flux_calculator=_lib.FLUX[flux],
            reconstruction=_lib.RECON[recon],
            limiter=_lib.LIMITER[limiter],
            max_lim_grad=float(maxLimGrad),
            k=0.0 if suth else float(tr.k),
            mu=0.0 if suth else float(tr.mu),
            diff=0.0 if suth else float(tr.diff),
            cfl=float(nst.parameters.get("cfl", 0.5)),
            conduction_stability_factor=float(nst.parameters.get("conductionStabilityFactor", 0.0)),
            viscous_stability_factor=float(nst.parameters.get("viscousStabilityFactor", 0.0)),
            ausmpup_minf=float(getattr(nst.fluxCalculator, "mInf", 0.0)),
            compact_species=int(compact),
            transport=int(suth),
            transport_disabled=tr.disabled_bits() if suth else 0,
        )
        if form == "mixed" and flux == "ausmpup":
            raise ValueError("the mixed form has no AusmpUp flux calculator: use form='fp64'")
        pgs_proc = [p for p in processes if isinstance(p, PressureGradientScaling)]
        flux_pgs = getattr(nst.fluxCalculator, "pgs", None)
        if len(pgs_proc) > 1:
            raise ValueError("at most one PressureGradientScaling process")
        if pgs_proc and flux == "ausmpup" and flux_pgs is not pgs_proc[0]:
            raise ValueError("AusmpUp must reference the PressureGradientScaling process of the solver (pgs: *pgs)")
        pgs = pgs_proc[0] if pgs_proc else flux_pgs
        if compact and nst.pgs is not None:
            if pgs is not None and pgs is not nst.pgs:
                raise ValueError("the compact process and AusmpUp must share one PressureGradientScaling")
            pgs = nst.pgs
        self.mesh = mesh
        self.phys = phys
        self.form = form
        self.n_species = n_species
        self.dim = mesh.dim
        self.n_comp = mesh.dim + 2 + n_species
        h = ctypes.c_void_p()
        fm = {"fp64": _lib.FORM_FP64, "mixed": _lib.FORM_MIXED}[form]
        _lib.check(_lib.lib().abl_fv_create(ctypes.byref(mesh.desc), ctypes.byref(phys), fm, ctypes.byref(h)), None,
                   "abl_fv_create")
        self._h = h
        self._theta_slots = []
        self._ghosts = []  # (Ghost, component slice, boundary faces, interior cells, ghost cells)
        self.pgs = pgs
        if pgs is not None:
            _lib.check(_lib.lib().abl_fv_set_pgs(h, ctypes.byref(pgs.desc())), h, "abl_fv_set_pgs")
            pgs._solver = weakref.ref(self)
        for bc in boundaryConditions:
            for slot in bc.labelIds:
                self.set_boundary(slot, bc)

    def __del__(self):
        if getattr(self, "_h", None):
            try:
                _lib.lib().abl_fv_destroy(self._h)
            except Exception:  # interpreter shutdown
                pass
            self._h = None

    @property
    def handle(self):
        return self._h

    def info(self) -> _lib.Info:
        inf = _lib.Info()
        _lib.check(_lib.lib().abl_fv_get_info(self._h, ctypes.byref(inf)), self._h, "abl_fv_get_info")
        return inf

    def set_boundary(self, slot: int, bc) -> None:
        desc = _lib.BcDesc()
        yi_buf = None
        if isinstance(bc, InflowGhost):
            desc.kind = _lib.BC_INFLOW
            desc.rho = float(bc.rho)
            vel = list(bc.velocity) + [0.0] * (3 - len(bc.velocity))
            desc.vel = (ctypes.c_double * 3)(*vel)
            desc.T = float(bc.T)
            desc.enforce_at_face = int(bool(bc.enforceAtFace))
            if bc.Yi is not None:
                if len(bc.Yi) != self.n_species:
                    raise ValueError(f"InflowGhost.Yi needs {self.n_species} mass fractions, got {len(bc.Yi)}")
                yi_buf = (ctypes.c_double * len(bc.Yi))(*bc.Yi)
                desc.Yi = ctypes.cast(yi_buf, ctypes.POINTER(ctypes.c_double))
            self._theta_slots.append(slot)
        elif isinstance(bc, OutflowGhost):
            desc.kind = _lib.BC_OUTFLOW
        elif isinstance(bc, Ghost):
            desc.kind = _lib.BC_GHOST
            self._register_ghost(slot, bc)
        else:
            raise TypeError(f"unsupported boundary condition {bc!r}")
        _lib.check(_lib.lib().abl_fv_set_boundary(self._h, int(slot), ctypes.byref(desc)), self._h,
                   "abl_fv_set_boundary")

    def _register_ghost(self, slot: int, bc: "Ghost") -> None:
        a = self.mesh.arrays
        if a.face_bc is None:
            raise ValueError("the mesh has no boundary faces")
        faces = np.nonzero(a.face_bc == slot)[0]
        ne = self.dim + 2
        comps = slice(0, ne) if bc.fieldName == "euler" else slice(ne, ne + self.n_species)
        for i, (b, c, f, _, _) in enumerate(self._ghosts):
            if b is bc:  # one update call per condition: the faces of all its label ids together
                f = np.concatenate([f, faces])
                self._ghosts[i] = (b, c, f, a.face_cells[f, 0].astype(np.int64), a.face_cells[f, 1].astype(np.int64))
                return
        self._ghosts.append((bc, comps, faces, a.face_cells[faces, 0].astype(np.int64),
                             a.face_cells[faces, 1].astype(np.int64)))

    def insert_boundary_values(self, time: float, U) -> None:
        """Writes the ghost rows of the Ghost conditions into the device state U (the reference's
        DMPlexInsertBoundaryValues before ComputeRHSFunction, timeStepper.cpp:347-349)."""
        if not self._ghosts:
            return
        import torch

        a = self.mesh.arrays
        for bc, comps, faces, interior, ghosts in self._ghosts:
            if faces.size == 0 or bc.updateFunction is None:
                continue
            xI = U[comps][:, torch.as_tensor(interior, device=U.device)].T.cpu().numpy()
            xG = np.asarray(bc.updateFunction(float(time), a.face_centroid[faces], a.face_normal[faces], xI),
                            dtype=np.float64)
            if xG.shape != xI.shape:
                raise ValueError(f"Ghost update function returned shape {xG.shape}, expected {xI.shape}")
            U[comps.start:comps.stop, torch.as_tensor(ghosts, device=U.device)] = torch.as_tensor(
                xG.T.copy(), device=U.device)

    # ---- hot path
    @staticmethod
    def _stream_ptr(stream):
        if stream is None:
            import torch

            return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        return ctypes.c_void_p(int(stream))

    def _check_tensor(self, t, name):
        import torch

        if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()):
            raise TypeError(f"{name} must be a contiguous float64 device tensor")
        if tuple(t.shape) != (self.n_comp, self.mesh.n_cells):
            raise ValueError(f"{name} must have shape {(self.n_comp, self.mesh.n_cells)}, got {tuple(t.shape)}")

    def compute_rhs(self, time: float, U, R, accumulate: bool = False, stream=None) -> None:
        self._check_tensor(U, "U")
        self._check_tensor(R, "R")
        self.insert_boundary_values(time, U)
        _lib.check(
            _lib.lib().abl_fv_rhs(self._h, float(time), ctypes.c_void_p(U.data_ptr()), ctypes.c_void_p(R.data_ptr()),
                                  int(bool(accumulate)), self._stream_ptr(stream)),
            self._h,
            "abl_fv_rhs",
        )

    def ComputeRHSFunction(self, time: float, locX, locF) -> int:
        """Reference semantics: locF is accumulated into (the caller zeroes it); returns 0 (PetscErrorCode)."""
        self.compute_rhs(time, locX, locF, accumulate=True)
        return 0

    def rhs_vjp(self, time: float, U, Rbar, Ubar, stream=None) -> np.ndarray:
        self._check_tensor(U, "U")
        self._check_tensor(Rbar, "Rbar")
        self._check_tensor(Ubar, "Ubar")
        self.insert_boundary_values(time, U)
        theta = np.zeros(3 * 16)
        _lib.check(
            _lib.lib().abl_fv_rhs_vjp(self._h, float(time), ctypes.c_void_p(U.data_ptr()),
                                      ctypes.c_void_p(Rbar.data_ptr()), ctypes.c_void_p(Ubar.data_ptr()),
                                      theta.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), self._stream_ptr(stream)),
            self._h,
            "abl_fv_rhs_vjp",
        )
        return theta.reshape(16, 3)

    def GetInterpolatedFaceVectors(self, U, stream=None):
        """FaceInterpolant::GetInterpolatedFaceVectors (faceInterpolant.cpp:109-249) on the device state U: returns
        (faceSolution [n_comp, F], faceSolutionGrad [n_comp, dim, F], faceAux [naux, F], faceAuxGrad [naux, dim, F])
        with aux = (T, velocity, Yi)."""
        import torch

        self._check_tensor(U, "U")
        F, nc, dim = self.mesh.n_faces, self.n_comp, self.dim
        na = 1 + dim + self.n_species
        out = [torch.empty((nc, F), dtype=torch.float64, device=U.device),
               torch.empty((nc, dim, F), dtype=torch.float64, device=U.device),
               torch.empty((na, F), dtype=torch.float64, device=U.device),
               torch.empty((na, dim, F), dtype=torch.float64, device=U.device)]
        _lib.check(_lib.lib().abl_fv_face_interpolate(self._h, ctypes.c_void_p(U.data_ptr()),
                                                      *[ctypes.c_void_p(t.data_ptr()) for t in out],
                                                      self._stream_ptr(stream)), self._h, "abl_fv_face_interpolate")
        return tuple(out)

    def set_strategy(self, strategy: str) -> None:
        """"auto" (the default at creation: tile or face_flux from the mesh's tile statistics), "march" (a workgroup
        walks a sequence of face-sharing tiles, computing the next tile's gradients while it evaluates and gathers
        the current one: the z-layers of a column-ordered box), "fused" (a workgroup per tile of cells computes the
        gradients of the tile and its halo, the tile's face fluxes and its residuals in one kernel), "tile" (a
        workgroup per tile of cells evaluates the tile's faces into LDS and gathers its cells), "face_flux" (stored
        once in HBM, gathered per cell) or "recompute" (each face flux evaluated by both cells)."""
        _lib.check(_lib.lib().abl_fv_set_strategy(self._h, _lib.STRATEGY[strategy]), self._h, "abl_fv_set_strategy")

    def has_strategy(self, strategy: str) -> bool:
        """Whether this handle offers a strategy (the fused and march ones need the FP64 form without species, no
        limiter or Barth-Jespersen and well-conditioned least-squares stencils).  The current strategy is kept."""
        current = self.strategy()
        ok = _lib.lib().abl_fv_set_strategy(self._h, _lib.STRATEGY[strategy]) == 0
        _lib.check(_lib.lib().abl_fv_set_strategy(self._h, _lib.STRATEGY[current]), self._h, "abl_fv_set_strategy")
        return ok

    def set_tile_size(self, cells: int = 0) -> None:
        """Cells per tile of the tile strategy (0 = default choice; 64, 128, 256 or 512)."""
        _lib.check(_lib.lib().abl_fv_set_tile_size(self._h, int(cells)), self._h, "abl_fv_set_tile_size")

    def strategy(self) -> str:
        """The strategy compute_rhs runs ("auto" resolved)."""
        v = ctypes.c_int32()
        _lib.check(_lib.lib().abl_fv_get_strategy(self._h, ctypes.byref(v)), self._h, "abl_fv_get_strategy")
        return {code: name for name, code in _lib.STRATEGY.items()}[v.value]

    def set_profiling(self, enable: bool) -> None:
        _lib.check(_lib.lib().abl_fv_set_profiling(self._h, int(bool(enable))), self._h, "abl_fv_set_profiling")

    def phase_times(self, reset: bool = True) -> dict:
        """Device milliseconds and launch counts per RHS phase (event-timed on the launch stream)."""
        n = len(_lib.PHASES)
        ms = (ctypes.c_double * n)()
        cnt = (ctypes.c_int64 * n)()
        _lib.check(_lib.lib().abl_fv_get_phase_times(self._h, ms, cnt, int(bool(reset))), self._h,
                   "abl_fv_get_phase_times")
        return {name: (ms[i], cnt[i]) for i, name in enumerate(_lib.PHASES)}

    def compute_cfl_dt(self, U, stream=None) -> float:
        """NavierStokesTransport::ComputeCflTimeStep on the device state U (host float; synchronises)."""
        self._check_tensor(U, "U")
        dt = ctypes.c_double()
        _lib.check(_lib.lib().abl_fv_cfl_dt(self._h, ctypes.c_void_p(U.data_ptr()), ctypes.byref(dt),
                                            self._stream_ptr(stream)), self._h, "abl_fv_cfl_dt")
        return float(dt.value)

    def ComputePhysicsTimeSteps(self, U, stream=None) -> dict:
        """FiniteVolumeSolver::ComputePhysicsTimeSteps (finiteVolumeSolver.cpp:381-393): the registered time-step
        functions' minima over the interior cells, by name ("cfl" with a flux calculator; "cond" / "visc" when
        NavierStokesTransport's conductionStabilityFactor / viscousStabilityFactor are positive)."""
        self._check_tensor(U, "U")
        dt = (ctypes.c_double * 3)()
        reg = ctypes.c_int32()
        _lib.check(_lib.lib().abl_fv_time_steps(self._h, ctypes.c_void_p(U.data_ptr()), dt, ctypes.byref(reg),
                                                self._stream_ptr(stream)), self._h, "abl_fv_time_steps")
        return {name: float(dt[i]) for i, name in enumerate(("cfl", "cond", "visc")) if reg.value >> i & 1}

    def ComputePhysicsTimeStep(self, U, stream=None) -> float:
        """FiniteVolumeSolver::ComputePhysicsTimeStep (finiteVolumeSolver.cpp:371-379): the smallest registered step
        (1e10 when none)."""
        self._check_tensor(U, "U")
        dt = ctypes.c_double()
        _lib.check(_lib.lib().abl_fv_physics_dt(self._h, ctypes.c_void_p(U.data_ptr()), ctypes.byref(dt),
                                                self._stream_ptr(stream)), self._h, "abl_fv_physics_dt")
        return float(dt.value)

    def axpy(self, dt: float, R, U, stream=None) -> None:
        """Forward-Euler stage on the interior rows: U += dt * R (FP64, on the device)."""
        self._check_tensor(U, "U")
        self._check_tensor(R, "R")
        _lib.check(_lib.lib().abl_fv_axpy(self._h, float(dt), ctypes.c_void_p(R.data_ptr()),
                                          ctypes.c_void_p(U.data_ptr()), self._stream_ptr(stream)), self._h,
                   "abl_fv_axpy")

    def ts_euler(self, U, max_time: float, t0: float = 0.0, dt0: float = 0.1, safety: float = 1.0,
                 max_steps: int = 10**9, R=None, stream=None):
        """Forward-Euler integration of U in place to max_time under the physics step-size adaptor (the smallest of
        the registered CFL / conduction / viscous steps) with matchstep (the reference TimeStepper with ts_adapt_type
        physics).  Returns (final time, steps)."""
        import torch

        if any(g[0].updateFunction is not None for g in self._ghosts):
            raise NotImplementedError("ts_euler runs on the device and cannot call the host Ghost update functions: "
                                      "step with compute_rhs and axpy instead")
        self._check_tensor(U, "U")
        R = torch.empty_like(U) if R is None else R
        self._check_tensor(R, "R")
        t = ctypes.c_double()
        n = ctypes.c_int64()
        _lib.check(_lib.lib().abl_fv_ts_euler(self._h, ctypes.c_void_p(U.data_ptr()), ctypes.c_void_p(R.data_ptr()),
                                              float(t0), float(max_time), float(dt0), float(safety), int(max_steps),
                                              ctypes.byref(t), ctypes.byref(n), self._stream_ptr(stream)),
                   self._h, "abl_fv_ts_euler")
        return float(t.value), int(n.value)

    def check_state(self, U, stream=None) -> int:
        self._check_tensor(U, "U")
        n = ctypes.c_int64()
        _lib.check(_lib.lib().abl_fv_check_state(self._h, ctypes.c_void_p(U.data_ptr()), ctypes.byref(n),
                                                 self._stream_ptr(stream)), self._h, "abl_fv_check_state")
        return int(n.value)
