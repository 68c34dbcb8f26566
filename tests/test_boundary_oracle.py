"""CPU checks of the boundary and input conventions: the oracle's ghost kinds (caller-filled rows untouched,
enforceAtFace = 2 state - interior, essentialGhost.cpp:15-21), seeded states that follow the physical cell under
every cell ordering, and the mixed form's refusal of AusmpUp before any device work."""
import numpy as np
import pytest

from ablate_amd import fv, synthetic
from ablate_amd.mesh import BoxMesh, TetBoxMesh
from oracle import binding as orc
from tests.fvcases import Case, ghost_and_at_face


def test_oracle_caller_ghost_and_at_face():
    mesh = BoxMesh([5, 4, 3], [0, 0, 0], [1, 1, 1], [False, True, True])
    case = Case("g", mesh, n_species=2, bcs=ghost_and_at_face(2))
    U0 = case.initial_state()
    a = mesh.arrays
    _, Uf = case.oracle_rhs(U0.copy())
    g0 = a.face_cells[a.face_bc == 0]
    np.testing.assert_array_equal(Uf[:, g0[:, 1]], U0[:, g0[:, 1]])  # caller rows kept
    f1 = a.face_cells[a.face_bc == 1]
    b = case.bcs[1]
    cv = 287.0 / (1.4 - 1.0)  # PerfectGas cv = R / (gamma - 1), perfectGas.cpp:128-133
    st = np.array([b["rho"], b["rho"] * (cv * b["T"] + 0.5 * np.dot(b["vel"], b["vel"])), *(b["rho"] * np.array(b["vel"])),
                   *(b["rho"] * np.array(b["Yi"]))])
    np.testing.assert_array_equal(Uf[:, f1[:, 1]], 2.0 * st[:, None] - U0[:, f1[:, 0]])
    # the face average of ghost and interior is the inflow state (what enforceAtFace is for)
    np.testing.assert_allclose(0.5 * (Uf[:, f1[:, 1]] + Uf[:, f1[:, 0]]), np.repeat(st[:, None], len(f1), 1),
                               rtol=1e-14)


def test_oracle_rejects_unset_slot():
    mesh = BoxMesh([4, 4, 4], [0, 0, 0], [1, 1, 1], [False, True, True])
    case = Case("u", mesh, bcs={0: {"kind": "outflow"}})
    with pytest.raises(RuntimeError, match="status"):
        orc.rhs(mesh.arrays, case.oracle_physics(), case.initial_state(), {0: {"kind": 2}})


@pytest.mark.parametrize("make", [
    lambda o: BoxMesh([6, 5, 4], [0, 0, 0], [1, 1, 1], [True] * 3, ordering=o),
    lambda o: TetBoxMesh(3, 0.2, ordering=o),
], ids=["box", "tets"])
def test_seeded_state_follows_the_physical_cell(make):
    """(seed, cell, comp) counters on the canonical cell id: every ordering of the same mesh sees the same state
    on the same cell (matched by centroid)."""
    first = make("natural")
    orders = ["brick", "rcm"] if isinstance(first, BoxMesh) else ["rcm", "morton"]
    meshes = [first] + [make(o) for o in orders]
    ref = None
    for m in meshes:
        U = synthetic.taylor_green(m, 0.3, n_species=3)
        NI = m.n_interior
        key = np.lexsort(np.round(m.arrays.cell_centroid[:NI], 12).T[::-1])
        S = U[:, :NI][:, key]
        if ref is None:
            ref = S
        else:
            np.testing.assert_array_equal(S, ref)


def test_mixed_ausmpup_is_refused_before_device_work():
    mesh = BoxMesh([4, 4, 4], [0, 0, 0], [1, 1, 1], [True] * 3)
    eos = fv.PerfectGas({"gamma": 1.4, "Rgas": 287.0})
    with pytest.raises(ValueError, match="AusmpUp"):
        fv.FiniteVolumeSolver(mesh, [fv.NavierStokesTransport({}, eos, fv.AusmpUp(0.1))], {}, form="mixed")
