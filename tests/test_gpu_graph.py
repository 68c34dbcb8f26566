"""abl_fv_rhs launches kernels only (no allocation, memset or host synchronisation), so a caller can capture it in a
HIP graph: every replay, with the caller's own eager work and synchronisations between replays, gives bitwise the
eager residual of the state the graph reads."""
import pytest

from ablate_amd.mesh import BoxMesh
from tests.fvcases import Case, inflow_x

pytestmark = pytest.mark.gpu


def graph_cases():
    box3 = BoxMesh([12, 12, 12], [0, 0, 0], [1, 1, 1], [True] * 3)
    chan = BoxMesh([8, 6, 6], [0, 0, 0], [1, 1, 1], [False, True, True])
    return [
        Case("c1_euler2d_first_order", BoxMesh([64, 64], [0, 0], [1, 1], [True, True]), state="random2d"),
        Case("ns_lsq_bj", box3, recon="leastsquares", limiter="barthjespersen", mu=1.8e-5, k=0.0251125),
        Case("species_inflow_ghost_fill", chan, n_species=3, recon="leastsquares", limiter="barthjespersen", mu=0.5,
             k=5.0, diff=0.3, bcs=inflow_x(3)),
    ]


CASES = graph_cases()


@pytest.mark.parametrize("strategy", ["tile", "face_flux"])
@pytest.mark.parametrize("case", CASES, ids=[c.name for c in CASES])
def test_rhs_graph_replays_match_eager(case, strategy, torch_cuda):
    torch = torch_cuda
    print(f"[graph] {case.name} {strategy}", flush=True)
    solver = case.solver()
    solver.set_strategy(strategy)
    U0 = torch.tensor(case.initial_state(), dtype=torch.float64, device="cuda")
    U = U0.clone()
    R = torch.zeros_like(U)
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        solver.compute_rhs(0.0, U, R)  # warm-up on the capture stream
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        solver.compute_rhs(0.0, U, R)
    for i in range(4):
        U.copy_(U0 * (1.0 + 0.01 * i))  # the graph reads U in place
        R.fill_(float("nan"))
        g.replay()
        torch.cuda.synchronize()
        Re = torch.zeros_like(R)
        solver.compute_rhs(0.0, U, Re)  # eager work between replays
        torch.cuda.synchronize()
        assert torch.equal(R, Re), f"replay {i}"
