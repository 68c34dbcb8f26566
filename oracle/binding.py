"""ctypes binding of the CPU oracle (oracle/fv_oracle.c).

TEST INFRASTRUCTURE ONLY: used by tests/, the smoke check and the bench's CPU-baseline leg as the checker.
"""
from __future__ import annotations

import ctypes
import subprocess
from ctypes import POINTER, c_double, c_int, c_int32, c_int64
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "liboracle_fv.so"

FLUX = {"none": 0, "ausm": 1, "average": 2, "off": 3, "rusanov": 4, "ausmpup": 5}
RECON = {"upwind": 0, "none": 0, "leastsquares": 1}
LIMITER = {"none": 0, "zero": 1, "minmod": 2, "vanleer": 3, "vanalbada": 4, "superbee": 5, "mc": 6, "sin": 7,
           "barthjespersen": 8}
LEFT, RIGHT, NA = 1, 2, 0
BC_NONE, BC_INFLOW, BC_OUTFLOW, BC_GHOST = 0, 1, 2, 3

_dp = POINTER(c_double)


class OrcMesh(ctypes.Structure):
    _fields_ = [
        ("dim", c_int),
        ("n_cells", c_int64),
        ("n_interior", c_int64),
        ("n_faces", c_int64),
        ("face_cells", POINTER(c_int32)),
        ("face_normal", _dp),
        ("face_centroid", _dp),
        ("face_shift", _dp),
        ("cell_centroid", _dp),
        ("cell_volume", _dp),
        ("face_bc", POINTER(c_int32)),
    ]


class OrcPhysics(ctypes.Structure):
    _fields_ = [
        ("gamma", c_double),
        ("Rgas", c_double),
        ("n_species", c_int),
        ("flux", c_int),
        ("recon", c_int),
        ("limiter", c_int),
        ("max_lim_grad", c_double),
        ("k", c_double),
        ("mu", c_double),
        ("diff", c_double),
        ("ausmpup_minf", c_double),
        ("pgs_alpha", c_double),
        ("compact", c_int),
        ("transport", c_int),
        ("transport_disabled", c_int),
    ]


class OrcBc(ctypes.Structure):
    _fields_ = [("kind", c_int), ("rho", c_double), ("vel", c_double * 3), ("T", c_double), ("Yi", _dp),
                ("enforce_at_face", c_int)]


_lib = None

# Planted-mistake variants of the oracle (ORC_PLANT in fv_oracle.c): each changes one formula so that a test can
# show that its pin rejects the mistake.
PLANTS = {"project_sign": 1, "rusanov_min_speed": 2, "bj_neighbor": 3, "minmod_as_mc": 4, "face_value_left": 5,
          "conduction_sign": 6, "sutherland_cell_t": 7}


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


def _open(path: Path) -> ctypes.CDLL:
    L = ctypes.CDLL(str(path))
    _declare(L)
    return L


class planted:
    """Context manager: inside it every oracle call goes to the variant with the named planted mistake."""

    def __init__(self, name: str):
        self.path = HERE / "_build" / f"liboracle_fv_plant{PLANTS[name]}.so"

    def __enter__(self):
        global _lib
        subprocess.run(["make", "-s", "-C", str(HERE), str(self.path.relative_to(HERE))], check=True)
        self._saved = lib()
        _lib = _open(self.path)
        return self

    def __exit__(self, *exc):
        global _lib
        _lib = self._saved
        return False


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        _lib = _open(LIB_PATH)
    return _lib


def _declare(L: ctypes.CDLL) -> None:
    """Argument and result types of the oracle's exported functions."""
    d = c_double
    for name in ("orc_ausm", "orc_average_flux", "orc_off_flux"):
        getattr(L, name).argtypes = [d] * 8 + [_dp, _dp]
        getattr(L, name).restype = c_int
    L.orc_ausmpup.argtypes = [d] * 10 + [_dp, _dp]
    L.orc_ausmpup.restype = c_int
    for name, args in {
        "orc_pg_pressure": [_dp, c_int, d, d],
        "orc_pg_temperature": [_dp, c_int, d, d],
        "orc_pg_internal_energy": [_dp, c_int],
        "orc_pg_internal_energy_T": [d, d, d],
        "orc_pg_speed_of_sound": [_dp, c_int, d, d],
        "orc_pg_speed_of_sound_T": [_dp, d, d, d],
        "orc_pg_pressure_T": [_dp, d, d],
        "orc_pg_sensible_enthalpy_T": [d, d, d],
        "orc_pg_cv": [d, d],
        "orc_pg_cp": [d, d],
        "orc_limiter": [c_int, d],
    }.items():
        getattr(L, name).argtypes = args
        getattr(L, name).restype = d
    L.orc_ns_advection_flux.argtypes = [c_int, _dp, _dp, _dp, d, d, c_int, _dp]
    L.orc_ns_advection_flux.restype = c_int
    L.orc_stress_tensor.argtypes = [c_int, d, _dp, _dp]
    L.orc_stress_tensor.restype = None
    L.orc_lsq_weights.argtypes = [c_int, c_int, _dp, _dp]
    L.orc_lsq_weights.restype = c_int
    L.orc_rhs.argtypes = [POINTER(OrcMesh), POINTER(OrcPhysics), POINTER(OrcBc), c_int, d, _dp, _dp]
    L.orc_rhs.restype = c_int
    L.orc_create.argtypes = [POINTER(OrcMesh), POINTER(OrcPhysics), POINTER(ctypes.c_void_p)]
    L.orc_create.restype = c_int
    L.orc_rhs_ctx.argtypes = [ctypes.c_void_p, POINTER(OrcBc), c_int, d, _dp, _dp]
    L.orc_rhs_ctx.restype = c_int
    L.orc_destroy.argtypes = [ctypes.c_void_p]
    L.orc_destroy.restype = None
    L.orc_cell_gradients.argtypes = [POINTER(OrcMesh), POINTER(OrcPhysics), _dp, _dp]
    L.orc_cell_gradients.restype = c_int
    L.orc_min_cell_radius.argtypes = [POINTER(OrcMesh), _dp]
    L.orc_min_cell_radius.restype = None
    L.orc_cfl_dt.argtypes = [POINTER(OrcMesh), POINTER(OrcPhysics), d, _dp, _dp]
    L.orc_cfl_dt.restype = d
    L.orc_pgs_update.argtypes = [POINTER(OrcMesh), POINTER(OrcPhysics), _dp, d, d, d, d, _dp, _dp]
    L.orc_pgs_update.restype = None
    L.orc_rusanov_flux.argtypes = [c_int, _dp, _dp, _dp, d, d, c_int, _dp]
    L.orc_rusanov_flux.restype = None
    L.orc_face_interpolate.argtypes = [POINTER(OrcMesh), POINTER(OrcPhysics), _dp, _dp, _dp, _dp, _dp]
    L.orc_face_interpolate.restype = c_int
    L.orc_sutherland.argtypes = [d, d, d, c_int, _dp]
    L.orc_sutherland.restype = None
    L.orc_diffusion_dt.argtypes = [POINTER(OrcMesh), POINTER(OrcPhysics), d, d, _dp, _dp, _dp]
    L.orc_diffusion_dt.restype = None


def _ptr(a):
    return a.ctypes.data_as(_dp) if a is not None else None


def _mesh_struct(m):
    keep = []

    def arr(x, dtype):
        if x is None:
            return None
        y = np.ascontiguousarray(x, dtype=dtype)
        keep.append(y)
        return y

    fc = arr(m.face_cells, np.int32)
    fn, fcen = arr(m.face_normal, np.float64), arr(m.face_centroid, np.float64)
    fsh = arr(m.face_shift, np.float64)
    cc, cv = arr(m.cell_centroid, np.float64), arr(m.cell_volume, np.float64)
    fbc = arr(m.face_bc, np.int32)
    s = OrcMesh(
        dim=m.dim, n_cells=m.n_cells, n_interior=m.n_interior, n_faces=m.n_faces,
        face_cells=fc.ctypes.data_as(POINTER(c_int32)), face_normal=_ptr(fn), face_centroid=_ptr(fcen),
        face_shift=_ptr(fsh), cell_centroid=_ptr(cc), cell_volume=_ptr(cv),
        face_bc=fbc.ctypes.data_as(POINTER(c_int32)) if fbc is not None else None,
    )
    return s, keep


TRANSPORT = {"constant": 0, "sutherland": 1}
TRANSPORT_BITS = {"conductivity": 1, "viscosity": 2, "diffusivity": 4}


def physics(gamma=1.4, Rgas=287.0, n_species=0, flux="ausm", recon="upwind", limiter="none", max_lim_grad=1e10,
            k=0.0, mu=0.0, diff=0.0, ausmpup_minf=0.0, pgs_alpha=1.0, compact=False, transport="constant",
            disabled=()) -> OrcPhysics:
    """``transport`` "constant" (k, mu, diff) or "sutherland" (k, mu, diff ignored; ``disabled`` names the
    properties left out of enabledProperties)."""
    return OrcPhysics(gamma=gamma, Rgas=Rgas, n_species=n_species, flux=FLUX[flux], recon=RECON[recon],
                      limiter=LIMITER[limiter], max_lim_grad=max_lim_grad, k=k, mu=mu, diff=diff,
                      ausmpup_minf=ausmpup_minf, pgs_alpha=pgs_alpha, compact=int(bool(compact)),
                      transport=TRANSPORT[transport],
                      transport_disabled=sum(TRANSPORT_BITS[x] for x in disabled))


def sutherland(T: float, rho: float, cp: float, disabled=()) -> tuple[float, float, float]:
    """transport::Sutherland: (conductivity, viscosity, diffusivity) at temperature T, density rho and cp."""
    out = np.zeros(3)
    lib().orc_sutherland(T, rho, cp, sum(TRANSPORT_BITS[x] for x in disabled), _ptr(out))
    return float(out[0]), float(out[1]), float(out[2])


def _bc_array(bcs):
    """Boundary slots {slot: dict(kind, rho, vel, T, Yi, enforce_at_face)} as the oracle's orc_bc array; returns
    (array or None, slot count, buffers the array points into)."""
    if not bcs:
        return None, 0, []
    n_bcs = max(bcs) + 1
    bc_arr = (OrcBc * n_bcs)()
    ykeep = []
    for slot, b in bcs.items():
        bc_arr[slot].kind = b["kind"]
        bc_arr[slot].rho = b.get("rho", 0.0)
        v = list(b.get("vel", [0.0, 0.0, 0.0])) + [0.0] * 3
        bc_arr[slot].vel = (c_double * 3)(*v[:3])
        bc_arr[slot].T = b.get("T", 0.0)
        bc_arr[slot].enforce_at_face = int(bool(b.get("enforce_at_face", False)))
        if b.get("Yi") is not None:
            y = np.ascontiguousarray(b["Yi"], dtype=np.float64)
            ykeep.append(y)
            bc_arr[slot].Yi = _ptr(y)
    return bc_arr, n_bcs, ykeep


def rhs(mesh_arrays, phys: OrcPhysics, U_soa: np.ndarray, bcs=None, time: float = 0.0):
    """Oracle RHS from a zeroed locF. Returns (R_soa, U_soa with ghost rows filled)."""
    nc, N = U_soa.shape
    X = np.ascontiguousarray(U_soa.T)  # AoS [N][nc]
    F = np.zeros_like(X)
    ms, keep = _mesh_struct(mesh_arrays)
    bc_arr, n_bcs, ykeep = _bc_array(bcs)
    rc = lib().orc_rhs(ctypes.byref(ms), ctypes.byref(phys), bc_arr, n_bcs, time, _ptr(X), _ptr(F))
    if rc:
        raise RuntimeError(f"orc_rhs failed with status {rc}")
    return np.ascontiguousarray(F.T), np.ascontiguousarray(X.T)


class Context:
    """Oracle setup (LSQ weights, face stencils) kept for repeated RHS evaluations (CPU-baseline timing)."""

    def __init__(self, mesh_arrays, phys: OrcPhysics):
        self._ms, self._keep = _mesh_struct(mesh_arrays)
        self._phys = phys
        self._ctx = ctypes.c_void_p()
        rc = lib().orc_create(ctypes.byref(self._ms), ctypes.byref(self._phys), ctypes.byref(self._ctx))
        if rc:
            raise RuntimeError(f"orc_create failed with status {rc}")

    def rhs_aos(self, X: np.ndarray, F: np.ndarray, bcs=None) -> None:
        """locF += RHS(locX) on AoS arrays [N][nc]; ``bcs`` as for rhs() (None on periodic meshes)."""
        bc_arr, n_bcs, _keep = _bc_array(bcs)
        rc = lib().orc_rhs_ctx(self._ctx, bc_arr, n_bcs, 0.0, _ptr(X), _ptr(F))
        if rc:
            raise RuntimeError(f"orc_rhs_ctx failed with status {rc}")

    def close(self):
        if self._ctx:
            lib().orc_destroy(self._ctx)
            self._ctx = ctypes.c_void_p()

    def __del__(self):
        self.close()


def cell_gradients(mesh_arrays, phys: OrcPhysics, U_soa: np.ndarray) -> np.ndarray:
    nc, N = U_soa.shape
    X = np.ascontiguousarray(U_soa.T)
    G = np.zeros((N, nc, mesh_arrays.dim))
    ms, keep = _mesh_struct(mesh_arrays)
    rc = lib().orc_cell_gradients(ctypes.byref(ms), ctypes.byref(phys), _ptr(X), _ptr(G))
    if rc:
        raise RuntimeError(f"orc_cell_gradients failed with status {rc}")
    return G


# ---- leaf functions
def ausm(uL, aL, rhoL, pL, uR, aR, rhoR, pR):
    m, p = c_double(), c_double()
    d = lib().orc_ausm(uL, aL, rhoL, pL, uR, aR, rhoR, pR, ctypes.byref(m), ctypes.byref(p))
    return d, m.value, p.value


def average_flux(uL, aL, rhoL, pL, uR, aR, rhoR, pR):
    m, p = c_double(), c_double()
    d = lib().orc_average_flux(uL, aL, rhoL, pL, uR, aR, rhoR, pR, ctypes.byref(m), ctypes.byref(p))
    return d, m.value, p.value


def ausmpup(uL, aL, rhoL, pL, uR, aR, rhoR, pR, mInf=0.0, pgs_alpha=1.0):
    m, p = c_double(), c_double()
    d = lib().orc_ausmpup(mInf, pgs_alpha, uL, aL, rhoL, pL, uR, aR, rhoR, pR, ctypes.byref(m), ctypes.byref(p))
    return d, m.value, p.value


def off_flux(*args):
    m, p = c_double(), c_double()
    d = lib().orc_off_flux(*args, ctypes.byref(m), ctypes.byref(p))
    return d, m.value, p.value


def ns_advection_flux(normal, xL, xR, gamma=1.4, R=287.0, flux="ausm"):
    dim = len(normal)
    n = np.ascontiguousarray(normal, dtype=np.float64)
    a = np.ascontiguousarray(xL, dtype=np.float64)
    b = np.ascontiguousarray(xR, dtype=np.float64)
    out = np.zeros(dim + 2)
    lib().orc_ns_advection_flux(dim, _ptr(n), _ptr(a), _ptr(b), gamma, R, FLUX[flux], _ptr(out))
    return out


def rusanov_flux(normal, uL, uR, gamma=1.4, R=287.0, n_species=0):
    """Build-defined Rusanov flux of one face (area-weighted normal): euler then species components."""
    dim = len(normal)
    n = np.ascontiguousarray(normal, dtype=np.float64)
    a = np.ascontiguousarray(uL, dtype=np.float64)
    b = np.ascontiguousarray(uR, dtype=np.float64)
    out = np.zeros(dim + 2 + n_species)
    lib().orc_rusanov_flux(dim, _ptr(n), _ptr(a), _ptr(b), gamma, R, n_species, _ptr(out))
    return out


def face_interpolate(mesh_arrays, phys: OrcPhysics, U_soa: np.ndarray):
    """FaceInterpolant::GetInterpolatedFaceVectors: (faceSol [F, nc], faceSolGrad [F, nc, dim], faceAux [F, naux],
    faceAuxGrad [F, naux, dim]) with aux = (T, velocity, Yi); ghost rows of U are used as given."""
    nc, N = U_soa.shape
    dim = mesh_arrays.dim
    F = mesh_arrays.n_faces
    na = 1 + dim + phys.n_species
    X = np.ascontiguousarray(U_soa.T)
    fs, fsg = np.zeros((F, nc)), np.zeros((F, nc, dim))
    fa, fag = np.zeros((F, na)), np.zeros((F, na, dim))
    ms, keep = _mesh_struct(mesh_arrays)
    rc = lib().orc_face_interpolate(ctypes.byref(ms), ctypes.byref(phys), _ptr(X), _ptr(fs), _ptr(fsg), _ptr(fa),
                                    _ptr(fag))
    if rc:
        raise RuntimeError(f"orc_face_interpolate failed with status {rc}")
    return fs, fsg, fa, fag


def stress_tensor(dim, mu, grad_vel):
    g = np.ascontiguousarray(grad_vel, dtype=np.float64)
    out = np.zeros(dim * dim)
    lib().orc_stress_tensor(dim, mu, _ptr(g), _ptr(out))
    return out


def lsq_weights(dx: np.ndarray) -> np.ndarray:
    dx = np.ascontiguousarray(dx, dtype=np.float64)
    n, dim = dx.shape
    out = np.zeros((n, dim))
    lib().orc_lsq_weights(n, dim, _ptr(dx), _ptr(out))
    return out


def limiter(name: str, f: float) -> float:
    return lib().orc_limiter(LIMITER[name], f)


def pg(prop: str, euler, gamma=1.4, R=287.0, T=None):
    u = np.ascontiguousarray(euler, dtype=np.float64)
    dim = len(u) - 2
    L = lib()
    if prop == "pressure":
        return L.orc_pg_pressure(_ptr(u), dim, gamma, R)
    if prop == "temperature":
        return L.orc_pg_temperature(_ptr(u), dim, gamma, R)
    if prop == "internal_energy":
        return L.orc_pg_internal_energy(_ptr(u), dim) if T is None else L.orc_pg_internal_energy_T(T, gamma, R)
    if prop == "speed_of_sound":
        return L.orc_pg_speed_of_sound(_ptr(u), dim, gamma, R) if T is None else L.orc_pg_speed_of_sound_T(_ptr(u), T, gamma, R)
    if prop == "sensible_enthalpy":
        if T is None:
            T = L.orc_pg_temperature(_ptr(u), dim, gamma, R)
        return L.orc_pg_sensible_enthalpy_T(T, gamma, R)
    if prop == "cv":
        return L.orc_pg_cv(gamma, R)
    if prop == "cp":
        return L.orc_pg_cp(gamma, R)
    if prop == "density":
        return float(u[0])
    raise KeyError(prop)


def min_cell_radius(mesh_arrays) -> np.ndarray:
    """MIN_CELL_RADIUS mesh characteristic per cell (finiteVolumeSolver.cpp:185-215)."""
    ms, keep = _mesh_struct(mesh_arrays)
    r = np.zeros(mesh_arrays.n_cells)
    lib().orc_min_cell_radius(ctypes.byref(ms), _ptr(r))
    return r


def cfl_dt(mesh_arrays, phys: OrcPhysics, cfl: float, r: np.ndarray, U_soa: np.ndarray) -> float:
    """NavierStokesTransport::ComputeCflTimeStep (navierStokesTransport.cpp:209-285) on a SoA state."""
    ms, keep = _mesh_struct(mesh_arrays)
    X = np.ascontiguousarray(U_soa.T)
    rr = np.ascontiguousarray(r, dtype=np.float64)
    return float(lib().orc_cfl_dt(ctypes.byref(ms), ctypes.byref(phys), cfl, _ptr(rr), _ptr(X)))


def diffusion_dt(mesh_arrays, phys: OrcPhysics, fcond: float, fvisc: float, r: np.ndarray,
                 U_soa: np.ndarray) -> tuple[float, float]:
    """NavierStokesTransport::ComputeConductionTimeStep / ComputeViscousDiffusionTimeStep
    (navierStokesTransport.cpp:287-450) on a SoA state: (conduction dt, viscous dt)."""
    ms, keep = _mesh_struct(mesh_arrays)
    X = np.ascontiguousarray(U_soa.T)
    rr = np.ascontiguousarray(r, dtype=np.float64)
    out = np.zeros(2)
    lib().orc_diffusion_dt(ctypes.byref(ms), ctypes.byref(phys), fcond, fvisc, _ptr(rr), _ptr(X), _ptr(out))
    return float(out[0]), float(out[1])


def pgs_update(mesh_arrays, phys: OrcPhysics, U_soa: np.ndarray, dt: float, alpha: float, domain_length: float,
               max_alpha_allowed: float = 0.0, max_delta_pressure_fac: float = 0.0):
    """PressureGradientScaling::UpdatePreconditioner (pressureGradientScaling.cpp:16-137) on a SoA state; the
    constructor defaults (:10-11) apply to max_alpha_allowed < 1 and max_delta_pressure_fac <= 0.
    Returns (alpha, max_mach)."""
    ms, keep = _mesh_struct(mesh_arrays)
    X = np.ascontiguousarray(U_soa.T)
    a, mm = c_double(alpha), c_double()
    maa = max_alpha_allowed if max_alpha_allowed >= 1.0 else 100.0
    fac = max_delta_pressure_fac if max_delta_pressure_fac > 0.0 else 0.05
    lib().orc_pgs_update(ctypes.byref(ms), ctypes.byref(phys), _ptr(X), dt, domain_length, maa, fac,
                         ctypes.byref(a), ctypes.byref(mm))
    return a.value, mm.value


def petsc_euler_solve(rhs_fn, dt_fn, U0: np.ndarray, max_time: float, interior: int, dt0: float = 0.1,
                      safety: float = 1.0, matchstepfac=(0.01, 2.0), max_steps: int = 1000000,
                      adapt_on_updated_state: bool = False):
    """Forward-Euler time loop of the reference's TimeStepper with ts_adapt_type physics and
    ts_exact_final_time matchstep (timeStepper.cpp:80-130, adaptPhysics.cpp:3-59), as PETSc runs it:
      - the first step is min(dt0, safety * dt(U0)) (AdaptPhysics::AdaptInitializer);
      - each step U += dt * rhs(t, U); the adaptor then picks the next step from the state BEFORE the update
        (TSStep_Euler evaluates the step size while the solution vector still holds U_n);
      - matchstep: a next step that would pass max_time is cut to the remaining time; one that would leave less
        than matchstepfac[1] steps is halved, and within (1 + matchstepfac[0]) of the remainder it is stretched.
    adapt_on_updated_state=True is a deliberately different protocol (step size from U_{n+1}) for tests that show
    the reference tables tell the two apart.  Returns (U, t, steps)."""
    U = U0.copy()
    t = 0.0
    dt = min(dt0, safety * dt_fn(U))
    steps = 0
    while t < max_time and steps < max_steps:
        R = rhs_fn(t, U)
        h = safety * dt_fn(U)
        U[:, :interior] = U[:, :interior] + dt * R[:, :interior]
        if adapt_on_updated_state:
            h = safety * dt_fn(U)
        t_end = t + dt
        remaining = max_time - t_end
        if t_end < max_time and t_end + h > max_time:
            h = remaining
        if t_end < max_time and t_end + h < max_time and h * matchstepfac[1] > remaining:
            h = remaining / 2
        if t_end < max_time and t_end + h < max_time and h * (1.0 + matchstepfac[0]) > remaining:
            h = remaining
        t, dt = t_end, h
        steps += 1
    return U, t, steps
