"""The GPU FP64 RHS on the reference's manufactured-solution convergence table
(compressibleFlowMmsSourceTests.cpp:656-895): each level's RHS through the C ABI, with the exact-solution Ghost
boundary as caller-filled ghost cells, equals the oracle's to 1e-12 per field, and the residual of RHS + source
over the central third converges at the table's rates (+/- 0.2)."""
import numpy as np
import pytest

from ablate_amd import fv
from tests import mms
from tests.fvcases import rel_err_per_field
from tests.test_oracle_mms import _mesh, mms_state, oracle_physics, rates
from oracle import binding as orc

pytestmark = pytest.mark.gpu


def _solver(row: mms.MmsRow, mesh):
    c = row.consts
    eos = fv.PerfectGas({"gamma": c.gamma, "Rgas": c.R})
    flux = {"average": fv.AverageFlux(), "ausm": fv.Ausm(), "ausmpup": fv.AusmpUp(row.minf)}[row.flux]
    tr = fv.Constant(k=c.k, mu=c.mu) if (c.k or c.mu) else None
    return fv.FiniteVolumeSolver(mesh, [fv.NavierStokesTransport({"cfl": 0.5}, eos, flux, tr)],
                                 {"petscfv_type": row.recon, "petsclimiter_type": row.limiter},
                                 [fv.Ghost("euler", "walls", list(range(2 * c.dim)), mms.ghost_function(c))])


@pytest.mark.parametrize("row", mms.ROWS, ids=[r.name.replace(" ", "_") for r in mms.ROWS])
def test_gpu_mms_rates_match_reference_table(row, torch_cuda):
    torch = torch_cuda
    c = row.consts

    def gpu_residual(row_, mesh):
        print(f"[mms-gpu] {row_.name} cells={mesh.n_interior}", flush=True)
        U0 = mms_state(row_, mesh)
        s = _solver(row_, mesh)
        U = torch.tensor(U0, dtype=torch.float64, device="cuda")
        U[:, mesh.n_interior:] = 0.0  # the Ghost update function refills the ghost rows before the RHS
        R = torch.zeros_like(U)
        s.compute_rhs(0.0, U, R)
        torch.cuda.synchronize()
        Rg = R.cpu().numpy()
        Rref, _ = orc.rhs(mesh.arrays, oracle_physics(row_), U0.copy(),
                          {s_: {"kind": orc.BC_GHOST} for s_ in range(2 * c.dim)})
        err = rel_err_per_field(Rg, Rref, mesh.n_interior)
        assert err.max() <= 1e-12, err
        NI = mesh.n_interior
        return Rg[:, :NI] + mms.source(c, mesh.arrays.cell_centroid[:NI, : c.dim])

    l2, li = rates(row, gpu_residual)
    print(f"[mms-gpu] {row.name}: L2 {np.round(l2, 2)} Linf {np.round(li, 2)}", flush=True)
    np.testing.assert_allclose(l2, row.l2, atol=mms.RATE_TOL)
    np.testing.assert_allclose(li, row.linf, atol=mms.RATE_TOL)
