"""The RHS on meshes read from the reference's own Hdf5Serializer files (tests/golden/2D_Circle.hdf5: quadrilaterals,
2D_Circle_Tri.hdf5: triangles, initializer.3D.hdf5: hexahedra) through abl_fv_mesh_dmplex, against the oracle on the
same ingested geometry: inflow on the first "Face Sets" value, outflow ghosts on the others."""
from pathlib import Path

import numpy as np
import pytest

from ablate_amd.mesh import Hdf5Mesh
from tests.fvcases import Case, rel_err_per_field

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"


def hdf5_cases():
    out = []
    for name, fname, ordering in (("circle_quad", "2D_Circle.hdf5", "morton"), ("circle_tri", "2D_Circle_Tri.hdf5", "natural"),
                                  ("init3d_hex", "initializer.3D.hdf5", "rcm")):
        m = Hdf5Mesh(GOLDEN / fname, ordering)
        dim = m.dim
        vals = sorted(m.labels)
        inflow = {"kind": "inflow", "rho": 1.2, "vel": [30.0, 5.0, 0.0][:dim], "T": 300.0, "Yi": [0.3, 0.7]}
        bcs = {v: ({"kind": "outflow"} if i else inflow) for i, v in enumerate(vals)}
        state = "random2d" if dim == 2 else "tg"
        # a shifted Taylor-Green field: unshifted, its symmetries on this 10 x 10 x 5 box make some least-squares
        # gradients pure round-off, and the limiter ratio 0.5 du / (grad . dx) of two round-off values is arbitrary
        phase = (0.1, 0.23, 0.37)
        out.append(Case(f"hdf5_{name}_lsq_bj_ns", m, recon="leastsquares", limiter="barthjespersen", mu=0.02,
                        k=0.05, state=state, phase=phase, bcs={s: dict(b, Yi=None) for s, b in bcs.items()}))
        out.append(Case(f"hdf5_{name}_species_ausm", m, flux="ausm", n_species=2, recon="leastsquares",
                        limiter="minmod", mu=0.02, k=0.05, diff=0.01, state=state, phase=phase, bcs=bcs))
    return out


CASES = hdf5_cases()


@pytest.mark.parametrize("strategy", ["auto", "tile", "face_flux", "recompute"])
@pytest.mark.parametrize("case", CASES, ids=[c.name for c in CASES])
def test_hdf5_rhs_matches_oracle(case, strategy, torch_cuda):
    torch = torch_cuda
    print(f"[hdf5] {case.name} {strategy} cells={case.mesh.n_interior}", flush=True)
    U0 = case.initial_state()
    Rref, _ = case.oracle_rhs(U0.copy())
    s = case.solver()
    s.set_strategy(strategy)
    U = torch.tensor(U0, dtype=torch.float64, device="cuda")
    R = torch.zeros_like(U)
    s.compute_rhs(0.0, U, R)
    torch.cuda.synchronize()
    err = rel_err_per_field(R.cpu().numpy(), Rref, case.mesh.n_interior)
    print(f"[hdf5] {case.name} {strategy} error {err}", flush=True)
    assert err.max() <= 1e-12
