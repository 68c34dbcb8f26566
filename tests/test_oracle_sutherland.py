"""transport::Sutherland (eos/transport/sutherland.cpp:9-122) in the oracle, pinned by:

  - the reference's table (sutherlandTests.cpp:141-145: T = 300, rho = 1.1, cp = 1001.1 -> k, mu, D within the
    tests' relative 1e-5), every property through the same law, and enabledProperties zeroing the others
    (sutherland.cpp:88-91);
  - the law's closed form at its reference temperature (mu(273) = 1.716e-5);
  - the face plumbing (NavierStokesTransport::DiffusionFlux and SpeciesTransport::DiffusionSpeciesFlux take the
    transport properties at the face's interpolated temperature and density, navierStokesTransport.cpp:464-468,
    speciesTransport.cpp:207-213): for linear temperature, velocity and mass fractions at unit density the face
    interpolant is exact (faceInterpolantTests.cpp:224-265), so the residual of every cell with a full set of faces
    is -(1/V) sum_f +/- F_f with F_f the diffusion flux of the exact face state and mu(T(x_f)); a planted variant
    that takes the left cell's temperature fails.
"""
import numpy as np
import pytest

from ablate_amd.mesh import BoxMesh
from oracle import binding as orc
from tests.test_oracle_pins import face_test_state, strip_ghosts

G, RG = 1.4, 287.0
CP = G * RG / (G - 1.0)


def test_reference_table():
    k, mu, d = orc.sutherland(300.0, 1.1, 1001.1)
    for got, want in ((k, 0.02615186), (mu, 1.8469051e-5), (d, 2.374829e-5)):
        assert abs((want - got) / want) < 1e-5


def test_enabled_properties_and_reference_temperature():
    k, mu, d = orc.sutherland(273.0, 2.0, CP)
    assert mu == pytest.approx(1.716e-5, rel=1e-15)
    assert k == pytest.approx(mu * CP / 0.707, rel=1e-15)
    assert d == pytest.approx(mu / 2.0 / 0.707, rel=1e-15)
    assert orc.sutherland(300.0, 1.1, CP, disabled=("viscosity",))[1] == 0.0
    assert orc.sutherland(300.0, 1.1, CP, disabled=("conductivity", "diffusivity"))[::2] == (0.0, 0.0)


def _suth_mu(T):
    r = T / 273.0
    return 1.716e-5 * np.sqrt(r) * r * (273.0 + 111.0) / (T + 111.0)


def _diffusion_deviation(dim, ns=2):
    """Largest deviation of the oracle residual from the exact-face-state diffusion flux divergence, relative to
    the largest residual, over the cells that have all 2 dim faces (a box without ghost cells)."""
    box = BoxMesh([5] * dim, [0.0] * dim, [1.0] * dim, [False] * dim, ordering="natural")
    a = strip_ghosts(box.arrays)
    x = a.cell_centroid[:, :dim]
    U, gaux = face_test_state(x, dim)  # rho = 1, T and velocity linear
    gY = np.array([[0.3, -0.2, 0.1], [-0.1, 0.25, 0.05]])[:ns, :dim]
    Y = 0.3 + x @ gY.T  # [n, ns]: linear mass fractions, rhoY = Y at rho = 1
    Us = np.vstack([U, Y.T])
    phys = orc.physics(n_species=ns, flux="none", transport="sutherland")
    R, _ = orc.rhs(a, phys, Us)
    # exact diffusion flux through each face from the exact face state
    xf = a.face_centroid[:, :dim]
    Uf, _ = face_test_state(xf, dim)
    cv = RG / (G - 1.0)
    T = (Uf[1] - 0.5 * np.sum(Uf[2:] ** 2, axis=0)) / cv
    mu = _suth_mu(T)
    k = mu * CP / 0.707
    gT, gV = gaux[0], gaux[1:]  # [dim], [dim][dim] (component, direction)
    div = np.trace(gV)
    tau = gV + gV.T - (2.0 / 3.0) * div * np.eye(dim)  # per unit mu
    n = a.face_normal[:, :dim]
    F = np.zeros((dim + 2 + ns, a.n_faces))
    F[2 : 2 + dim] = -(mu[None, :] * (tau @ n.T))
    vel = Uf[2:]
    F[1] = -np.einsum("fd,df->f", n, mu[None, :] * (tau @ vel)) - k * (n @ gT)
    F[2 + dim :] = -(mu / 0.707)[None, :] * (gY @ n.T)  # rho D = mu / Sc at rho = 1
    Rex = np.zeros_like(R)
    L, Rc = a.face_cells[:, 0], a.face_cells[:, 1]
    for p in range(F.shape[0]):
        np.add.at(Rex[p], L, -F[p] / a.cell_volume[L])
        np.add.at(Rex[p], Rc, F[p] / a.cell_volume[Rc])
    full = np.bincount(np.concatenate([L, Rc]), minlength=a.n_cells) == 2 * dim
    scale = np.max(np.abs(R[:, full]), axis=1)
    keep = scale > 0
    return np.max(np.abs(R[:, full] - Rex[:, full])[keep], axis=1) / scale[keep]


@pytest.mark.parametrize("dim", [2, 3])
def test_face_properties_at_face_temperature(dim):
    dev = _diffusion_deviation(dim)
    assert np.all(dev < 1e-9), dev


def test_face_temperature_planted_mistake_fails():
    with orc.planted("sutherland_cell_t"):
        dev = _diffusion_deviation(3)
    assert np.max(dev) > 1e-4, dev
