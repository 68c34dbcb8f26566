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


def test_noalloc_vjp_graph_replays_match_eager(torch_cuda):
    """abl_fv_rhs_vjp_noalloc with theta_bar = NULL launches kernels only (ghost fill, ghost-row fill, the cell-centric
    reverse): captured once, every replay gives bitwise the eager Ubar of the state and seed the graph reads."""
    import ctypes

    import numpy as np

    from ablate_amd import _lib

    torch = torch_cuda
    lib = _lib.lib()
    chan = BoxMesh([8, 6, 6], [0, 0, 0], [1, 1, 1], [False, True, True])
    case = Case("noalloc_graph", chan, n_species=3, bcs=inflow_x(3))
    solver = case.solver()
    U0 = torch.tensor(case.initial_state(), dtype=torch.float64, device="cuda")
    rng = np.random.default_rng(3)
    Rb0 = torch.tensor(rng.normal(size=tuple(U0.shape)) * (np.arange(U0.shape[1]) < chan.n_interior), device="cuda")
    U, Rb, Ub = U0.clone(), Rb0.clone(), torch.zeros_like(U0)
    ptr = lambda t: ctypes.c_void_p(t.data_ptr())

    def call(stream_ptr, ub):
        return lib.abl_fv_rhs_vjp_noalloc(solver.handle, 0.0, ptr(U), ptr(Rb), ptr(ub), None,
                                          ctypes.c_void_p(stream_ptr))

    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        assert call(side.cuda_stream, Ub) == 0, lib.abl_fv_last_error(solver.handle)  # warm-up on the capture stream
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        rc = call(torch.cuda.current_stream().cuda_stream, Ub)
    assert rc == 0, lib.abl_fv_last_error(solver.handle)
    for i in range(4):
        U.copy_(U0 * (1.0 + 0.01 * i))  # the graph reads U and Rbar in place
        Rb.copy_(Rb0 * (1.0 - 0.1 * i))
        Ub.fill_(float("nan"))
        g.replay()
        torch.cuda.synchronize()
        Ue = torch.full_like(Ub, float("nan"))
        assert call(torch.cuda.current_stream().cuda_stream, Ue) == 0  # eager work between replays
        torch.cuda.synchronize()
        assert torch.equal(Ub, Ue), f"replay {i}"
