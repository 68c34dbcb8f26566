"""The oracle's first-order RHS, CFL step size and the reference time-step protocol reproduce the reference's
shock-tube tables (compressibleShockTubeTests.cpp:209-314) to their 1e-6 tolerance."""
import numpy as np
import pytest

from oracle import binding as orc
from tests import shock_tube as st

CASES = st.cases()


def _oracle_run(case, **kw):
    mesh, U0, bcs = st.setup(case)
    a = mesh.arrays
    g = case["initialConditions"]["gamma"]
    phys = orc.physics(g, st.RGAS, 0, case["fluxCalculator"].lower(), "upwind", "none")
    kinds = {"inflow": 1, "outflow": 2}
    obcs = {s: dict(b, kind=kinds[b["kind"]]) for s, b in bcs.items()}
    r = orc.min_cell_radius(a)

    def rhs_fn(t, U):
        return orc.rhs(a, phys, U.copy(), obcs, t)[0]

    def dt_fn(U):
        return orc.cfl_dt(a, phys, case["cfl"], r, U)

    U, t, steps = orc.petsc_euler_solve(rhs_fn, dt_fn, U0, case["maxTime"], a.n_interior, **kw)
    return mesh, U, t, steps


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_oracle_reproduces_reference_shock_tube(case):
    mesh, U, t, steps = _oracle_run(case)
    assert t == pytest.approx(case["maxTime"], abs=1e-15)
    err = st.max_abs_error(case, st.primitives(mesh, U))
    assert max(err.values()) <= case["tolerance"], (steps, err)


def test_tables_distinguish_the_step_size_protocol():
    """Taking the step size from the updated state instead of U_n misses the Sod table by far more than 1e-6."""
    case = CASES[0]
    mesh, U, t, steps = _oracle_run(case, adapt_on_updated_state=True)
    err = st.max_abs_error(case, st.primitives(mesh, U))
    assert max(err.values()) > 1e-4, err


def test_min_cell_radius_uniform():
    case = CASES[0]
    mesh, _, _ = st.setup(case)
    r = orc.min_cell_radius(mesh.arrays)[: mesh.n_interior]
    np.testing.assert_allclose(r, 0.005, rtol=1e-13)
