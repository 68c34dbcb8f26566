"""transport::Sutherland on the GPU FP64 path beyond the oracle-parity cases: the properties follow the face
temperature (a state of uniform temperature and density gives the residual of the Constant model at mu(T_0),
k = mu(T_0) cp / Pr, D = mu(T_0) / (rho_0 Sc)), enabledProperties, and the forms / entry points that refuse it."""
import numpy as np
import pytest

from ablate_amd import fv
from ablate_amd._lib import NativeError
from ablate_amd.mesh import BoxMesh
from tests.fvcases import Case, inflow_x, rel_err_per_field

pytestmark = pytest.mark.gpu

G, RG = 1.4, 287.0
CP = G * RG / (G - 1.0)


def suth_mu(T):
    r = T / 273.0
    return 1.716e-5 * np.sqrt(r) * r * 384.0 / (T + 111.0)


def uniform_T_state(case, T0=350.0, rho0=1.3):
    """Seeded velocities and mass fractions at uniform density and temperature (so that every face has T_0)."""
    U = case.initial_state()
    ne = case.mesh.dim + 2
    vel = U[2:ne] / U[0]
    cv = RG / (G - 1.0)
    U[0] = rho0
    U[2:ne] = rho0 * vel
    U[1] = rho0 * (cv * T0 + 0.5 * np.sum(vel * vel, axis=0))
    if case.n_species:
        Y = U[ne:] / np.maximum(U[ne:].sum(axis=0), 1e-300)
        U[ne:] = rho0 * Y
    return U


def _rhs(solver, U0, torch):
    U = torch.tensor(U0, dtype=torch.float64, device="cuda")
    R = torch.zeros_like(U)
    solver.compute_rhs(0.0, U, R)
    torch.cuda.synchronize()
    return R.cpu().numpy()


@pytest.mark.parametrize("strategy", ["face_flux", "tile"])
def test_uniform_temperature_equals_constant_model(strategy, torch_cuda):
    torch = torch_cuda
    mesh = BoxMesh([6, 6, 6], [0, 0, 0], [1, 1, 1], [True, True, True], ordering="natural")
    T0, rho0 = 350.0, 1.3
    base = dict(n_species=2, recon="leastsquares", limiter="barthjespersen", flux="none")
    suth = Case("suth", mesh, transport="sutherland", **base)
    mu0 = suth_mu(T0)
    const = Case("const", mesh, mu=mu0, k=mu0 * CP / 0.707, diff=mu0 / rho0 / 0.707, **base)
    U0 = uniform_T_state(suth, T0, rho0)
    print(f"[sutherland] uniform T {strategy}", flush=True)
    s1, s2 = suth.solver(), const.solver()
    s1.set_strategy(strategy)
    s2.set_strategy(strategy)
    R1, R2 = _rhs(s1, U0, torch), _rhs(s2, U0, torch)
    err = rel_err_per_field(R1, R2, mesh.n_interior)
    assert np.all(err[1:] < 1e-12), err  # the density row has no diffusion flux


def test_enabled_properties(torch_cuda):
    """Only Viscosity enabled: the energy residual loses the conduction term, the species rows their diffusion."""
    torch = torch_cuda
    mesh = BoxMesh([6, 6, 6], [0, 0, 0], [1, 1, 1], [False, True, True])
    full = Case("full", mesh, flux="none", n_species=2, transport="sutherland", bcs=inflow_x(2))
    visc = Case("visc", mesh, flux="none", n_species=2, transport="sutherland",
                disabled=("conductivity", "diffusivity"), bcs=inflow_x(2))
    U0 = full.initial_state()
    Rf, Rv = _rhs(full.solver(), U0, torch), _rhs(visc.solver(), U0, torch)
    Rv_ref, _ = visc.oracle_rhs(U0.copy())
    assert rel_err_per_field(Rv, Rv_ref, mesh.n_interior).max() <= 1e-12
    ne = mesh.dim + 2
    np.testing.assert_array_equal(Rv[ne:, : mesh.n_interior], 0.0)
    np.testing.assert_array_equal(Rv[2:ne], Rf[2:ne])  # momentum: viscosity only in both
    assert np.abs(Rv[1] - Rf[1]).max() > 0.0


def test_refused_forms_and_adjoint(torch_cuda):
    torch = torch_cuda
    mesh = BoxMesh([4, 4, 4], [0, 0, 0], [1, 1, 1], [True, True, True])
    case = Case("suth", mesh, recon="leastsquares", limiter="barthjespersen", transport="sutherland")
    with pytest.raises(ValueError, match="Sutherland"):
        case.solver(form="mixed")
    solver = case.solver()
    U = torch.tensor(case.initial_state(), dtype=torch.float64, device="cuda")
    with pytest.raises(NativeError, match="Sutherland"):
        solver.rhs_vjp(0.0, U, torch.ones_like(U), torch.empty_like(U))
    with pytest.raises(ValueError, match="unknown transport"):
        fv.Sutherland(None, ["Enthalpy"]).disabled_bits()
