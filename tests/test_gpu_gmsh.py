"""The RHS on the reference's own Gmsh meshes (tests/golden/pipeFlow.msh: quadrilaterals, mixedCells2D.msh:
triangles + quadrilaterals) against the oracle on the same ingested geometry: inflow at the inlet label, outflow
ghosts at the outlet and walls."""
from pathlib import Path

import numpy as np
import pytest

from ablate_amd.mesh import GmshMesh
from tests.fvcases import Case, rel_err_per_field

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"


def gmsh_cases():
    out = []
    for name, fname, ordering in (("pipe", "pipeFlow.msh", "natural"), ("pipe_rcm", "pipeFlow.msh", "rcm"),
                                  ("mixed_cells", "mixedCells2D.msh", "natural")):
        m = GmshMesh(GOLDEN / fname, ordering)
        lab = m.labels
        inflow = {"kind": "inflow", "rho": 1.2, "vel": [30.0, 0.0], "T": 300.0, "Yi": [0.3, 0.7]}
        bcs = {lab["inlet"]: inflow, lab["outlet"]: {"kind": "outflow"}, lab["wall"]: {"kind": "outflow"}}
        out.append(Case(f"gmsh_{name}_lsq_bj_ns", m, recon="leastsquares", limiter="barthjespersen", mu=0.02,
                        k=0.05, state="random2d", bcs={s: dict(b, Yi=None) for s, b in bcs.items()}))
        out.append(Case(f"gmsh_{name}_species_rusanov", m, flux="rusanov", n_species=2, recon="leastsquares",
                        limiter="minmod", mu=0.02, k=0.05, diff=0.01, state="random2d", bcs=bcs))
    return out


CASES = gmsh_cases()


@pytest.mark.parametrize("strategy", ["tile", "face_flux", "recompute"])
@pytest.mark.parametrize("case", CASES, ids=[c.name for c in CASES])
def test_gmsh_rhs_matches_oracle(case, strategy, torch_cuda):
    torch = torch_cuda
    print(f"[gmsh] {case.name} {strategy} cells={case.mesh.n_interior}", flush=True)
    U0 = case.initial_state()
    Rref, _ = case.oracle_rhs(U0.copy())
    s = case.solver()
    s.set_strategy(strategy)
    U = torch.tensor(U0, dtype=torch.float64, device="cuda")
    R = torch.zeros_like(U)
    s.compute_rhs(0.0, U, R)
    torch.cuda.synchronize()
    err = rel_err_per_field(R.cpu().numpy(), Rref, case.mesh.n_interior)
    print(f"[gmsh] {case.name} error {err}", flush=True)
    assert err.max() <= 1e-12


@pytest.mark.parametrize("case", CASES[:2] + CASES[4:], ids=[c.name for c in CASES[:2] + CASES[4:]])
def test_gmsh_mixed_form(case, torch_cuda):
    torch = torch_cuda
    U0 = case.initial_state()
    Rref, _ = case.oracle_rhs(U0.copy())
    s = case.solver(form="mixed")
    U = torch.tensor(U0, dtype=torch.float64, device="cuda")
    R = torch.zeros_like(U)
    s.compute_rhs(0.0, U, R)
    torch.cuda.synchronize()
    err = rel_err_per_field(R.cpu().numpy(), Rref, case.mesh.n_interior)
    print(f"[gmsh] mixed {case.name} error {err}", flush=True)
    assert np.all(np.isfinite(R.cpu().numpy())) and err.max() <= 5e-6
