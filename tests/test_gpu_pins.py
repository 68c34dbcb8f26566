"""The GPU path on the mathematics-fixed pins of the second-order RHS: exactness for linear states (least-squares
gradients, ProjectToFace, face orientation and scatter reproduce the physical flux at the face centroids) and the
face interpolant's exact face values and gradients of linear fields (faceInterpolantTests.cpp:224-265, 1e-8),
through the C ABI, next to the oracle on the same inputs (1e-12)."""
import numpy as np
import pytest

from ablate_amd import fv
from ablate_amd.mesh import ArrayMesh, BoxMesh, TetBoxMesh
from oracle import binding as orc
from tests.test_oracle_pins import (LINEAR_CASES, LINEAR_MESHES, face_test_state, linear_deviation, linear_state,
                                    strip_ghosts)

pytestmark = pytest.mark.gpu


def _solver(mesh, flux, recon, limiter, mu=0.0, k=0.0, ghost=True):
    eos = fv.PerfectGas({"gamma": 1.4, "Rgas": 287.0})
    fc = {"ausm": fv.Ausm, "average": fv.AverageFlux, "rusanov": fv.Rusanov}[flux]()
    tr = fv.Constant(k=k, mu=mu) if (k or mu) else None
    bcs = [fv.Ghost("euler", "walls", list(range(2 * mesh.dim)))] if ghost else []  # rows written by the caller
    return fv.FiniteVolumeSolver(mesh, [fv.NavierStokesTransport({}, eos, fc, tr)],
                                 {"petscfv_type": recon, "petsclimiter_type": limiter}, bcs)


@pytest.mark.parametrize("strategy", ["recompute", "face_flux", "tile"])
@pytest.mark.parametrize("mesh_kind,flux,limiter", LINEAR_CASES)
def test_gpu_second_order_rhs_is_exact_for_linear_states(mesh_kind, flux, limiter, strategy, torch_cuda):
    torch = torch_cuda
    print(f"[pins] linear {mesh_kind} {flux} {limiter} {strategy}", flush=True)
    mesh = LINEAR_MESHES[mesh_kind]()
    U0 = linear_state(mesh, mesh.dim, np.random.default_rng(3))  # exact values in the ghost rows as well
    s = _solver(mesh, flux, "leastsquares", limiter)
    s.set_strategy(strategy)
    U = torch.tensor(U0, dtype=torch.float64, device="cuda")
    R = torch.zeros_like(U)
    s.compute_rhs(0.0, U, R)
    torch.cuda.synchronize()
    Rg = R.cpu().numpy()
    dev, n = linear_deviation(mesh, U0, Rg)
    assert n > 0 and dev < 1e-10, dev
    Ro, _ = orc.rhs(mesh.arrays, orc.physics(flux=flux, recon="leastsquares", limiter=limiter), U0.copy(),
                    {s_: {"kind": orc.BC_GHOST} for s_ in range(2 * mesh.dim)})
    NI = mesh.n_interior
    for p in range(U0.shape[0]):
        assert np.linalg.norm(Rg[p, :NI] - Ro[p, :NI]) <= 1e-12 * np.linalg.norm(Ro[p, :NI])


def _gpu_face_vectors(mesh, U0, torch):
    s = _solver(mesh, "ausm", "upwind", "none", mu=1.0, k=1.0, ghost=mesh.arrays.face_bc is not None)
    U = torch.tensor(U0, dtype=torch.float64, device="cuda")
    out = s.GetInterpolatedFaceVectors(U)
    torch.cuda.synchronize()
    return [t.cpu().numpy() for t in out]


@pytest.mark.parametrize("dim", [2, 3])
def test_gpu_face_interpolant_exact_for_linear_fields(dim, torch_cuda):
    """5^d box without ghost cells (faceInterpolantTests.cpp:63-73): every face value and gradient of the linear
    fields exact to 1e-8, and the same numbers as the oracle's stencils to 1e-12."""
    torch = torch_cuda
    box = BoxMesh([5] * dim, [0.0] * dim, [1.0] * dim, [False] * dim, ordering="natural")
    mesh = ArrayMesh(strip_ghosts(box.arrays))
    a = mesh.arrays
    U0, gaux = face_test_state(a.cell_centroid[:, :dim], dim)
    fs, fsg, fa, fag = _gpu_face_vectors(mesh, U0, torch)
    Ue, _ = face_test_state(a.face_centroid[:, :dim], dim)
    for p in [0] + list(range(2, dim + 2)):
        np.testing.assert_allclose(fs[p], Ue[p], atol=1e-8)
    np.testing.assert_allclose(fsg[2:], np.broadcast_to(gaux[1:, :, None], fsg[2:].shape), atol=1e-8)
    np.testing.assert_allclose(fag, np.broadcast_to(gaux[:, :, None], fag.shape), atol=1e-8)
    os_, osg, oa, oag = orc.face_interpolate(a, orc.physics(), U0)
    np.testing.assert_allclose(fs, os_.T, rtol=1e-12, atol=1e-12 * np.abs(os_).max())
    np.testing.assert_allclose(fsg, osg.transpose(1, 2, 0), rtol=1e-11, atol=1e-12 * np.abs(osg).max())
    np.testing.assert_allclose(fa, oa.T, rtol=1e-12, atol=1e-12 * np.abs(oa).max())
    np.testing.assert_allclose(fag, oag.transpose(1, 2, 0), rtol=1e-11, atol=1e-12 * np.abs(oag).max())


@pytest.mark.parametrize("make", [lambda: TetBoxMesh(3, 0.2, ordering="rcm"),
                                  lambda: BoxMesh([5, 4, 3], [0, 0, 0], [1, 2, 3], [False, True, False])],
                         ids=["tets", "hex_ghosts"])
def test_gpu_face_interpolant_matches_oracle_with_ghost_cells(make, torch_cuda):
    """With ghost cells (caller-filled rows: boundary faces included) the GPU face vectors equal the oracle's."""
    torch = torch_cuda
    from tests.fvcases import Case

    mesh = make()
    U0 = Case("fi", mesh, mach=0.4).initial_state()
    a = mesh.arrays
    g = a.face_cells[a.face_bc >= 0]
    U0[:, g[:, 1]] = U0[:, g[:, 0]] * 1.01  # ghost rows as a caller would write them
    fs, fsg, fa, fag = _gpu_face_vectors(mesh, U0, torch)
    os_, osg, oa, oag = orc.face_interpolate(a, orc.physics(), U0)
    for got, want in ((fs, os_.T), (fsg, osg.transpose(1, 2, 0)), (fa, oa.T), (fag, oag.transpose(1, 2, 0))):
        scale = np.abs(want).max(axis=-1, keepdims=True)
        assert np.abs(got - want).max(axis=-1).max() <= 1e-11 * scale.max(), np.abs(got - want).max()
