"""GPU parity of the FP64 RHS against the CPU oracle (norm-wise relative error per field <= 1e-12), through the
C ABI, plus determinism, accumulate semantics and AoS packing."""
import numpy as np
import pytest

from ablate_amd.mesh import BoxMesh
from tests.fvcases import Case, all_outflow, inflow_x, rel_err_per_field, standard_cases

pytestmark = pytest.mark.gpu

TOL_FP64 = 1e-12
CASES = standard_cases()
STRATEGIES = ["recompute", "face_flux", "tile", "fused", "march"]


def use_strategy(solver, strategy):
    """Selects a strategy; the fused and march ones exist only for some handles (FP64, no species, full-rank
    stencils)."""
    if strategy in ("fused", "march") and not solver.has_strategy(strategy):
        pytest.skip(f"{strategy} strategy unavailable for this handle")
    solver.set_strategy(strategy)


@pytest.mark.parametrize("strategy", STRATEGIES)
@pytest.mark.parametrize("case", CASES, ids=[c.name for c in CASES])
def test_rhs_matches_oracle(case, strategy, torch_cuda):
    torch = torch_cuda
    print(f"[parity] {case.name} {strategy} cells={case.mesh.n_interior} faces={case.mesh.n_faces}", flush=True)
    U0 = case.initial_state()
    Rref, Ufilled = case.oracle_rhs(U0.copy())
    solver = case.solver()
    use_strategy(solver, strategy)
    U = torch.tensor(U0, dtype=torch.float64, device="cuda")
    R = torch.zeros_like(U)
    solver.compute_rhs(0.0, U, R)
    torch.cuda.synchronize()
    Rg = R.cpu().numpy()
    err = rel_err_per_field(Rg, Rref, case.mesh.n_interior)
    assert np.all(np.isfinite(Rg)), case.name
    assert err.max() <= TOL_FP64, f"{case.name}: per-field relative error {err}"
    # ghost rows filled exactly like the oracle
    np.testing.assert_array_equal(U.cpu().numpy()[:, case.mesh.n_interior:], Ufilled[:, case.mesh.n_interior:])
    # ghost rows of R are zero with accumulate = 0
    assert not np.any(Rg[:, case.mesh.n_interior:])


@pytest.mark.parametrize("case", [c for c in CASES if c.name in ("c2_ns_lsq_bj", "c5_tets_rusanov", "c4_species_inflow")],
                         ids=lambda c: c.name)
@pytest.mark.parametrize("strategy", STRATEGIES)
def test_rhs_deterministic_and_accumulates(case, strategy, torch_cuda):
    torch = torch_cuda
    solver = case.solver()
    use_strategy(solver, strategy)
    U = torch.tensor(case.initial_state(), dtype=torch.float64, device="cuda")
    R1 = torch.zeros_like(U)
    R2 = torch.zeros_like(U)
    solver.compute_rhs(0.0, U, R1)
    solver.compute_rhs(0.0, U, R2)
    torch.cuda.synchronize()
    assert torch.equal(R1, R2), "two evaluations differ bitwise"
    base = torch.randn_like(U)
    R3 = base.clone()
    solver.ComputeRHSFunction(0.0, U, R3)  # reference semantics: adds into locF
    torch.cuda.synchronize()
    assert torch.equal(R3[:, : case.mesh.n_interior], base[:, : case.mesh.n_interior] + R1[:, : case.mesh.n_interior])
    assert torch.equal(R3[:, case.mesh.n_interior :], base[:, case.mesh.n_interior :])


@pytest.mark.parametrize("strategy", STRATEGIES)
def test_conservation_periodic(strategy, torch_cuda):
    """sum_c V_c R_c = 0 on a periodic mesh (every face flux enters two cells with opposite signs)."""
    torch = torch_cuda
    case = next(c for c in CASES if c.name == "c2_ns_lsq_bj")
    solver = case.solver()
    use_strategy(solver, strategy)
    U = torch.tensor(case.initial_state(), dtype=torch.float64, device="cuda")
    R = torch.zeros_like(U)
    solver.compute_rhs(0.0, U, R)
    vol = torch.tensor(case.mesh.arrays.cell_volume, device="cuda")
    tot = (R * vol).sum(dim=1)
    scale = (R.abs() * vol).sum(dim=1)
    assert torch.all(tot.abs() <= 1e-13 * scale), (tot, scale)


def test_strategies_agree_bitwise_per_face(torch_cuda):
    """Both strategies evaluate every face flux with the same arithmetic; the per-cell sums may differ only in
    summation association, so they agree to rounding."""
    torch = torch_cuda
    case = next(c for c in CASES if c.name == "c5_tets_rusanov")
    solver = case.solver()
    U = torch.tensor(case.initial_state(), dtype=torch.float64, device="cuda")
    Ra, Rb = torch.zeros_like(U), torch.zeros_like(U)
    solver.set_strategy("recompute")
    solver.compute_rhs(0.0, U, Ra)
    solver.set_strategy("face_flux")
    solver.compute_rhs(0.0, U, Rb)
    torch.cuda.synchronize()
    scale = Ra.abs().max(dim=1, keepdim=True).values
    assert torch.all((Ra - Rb).abs() <= 1e-14 * scale)


@pytest.mark.parametrize("case", CASES, ids=[c.name for c in CASES])
def test_tile_equals_face_flux_bitwise(case, torch_cuda):
    """The tile strategy evaluates the same face fluxes (faces on tile borders twice) and gathers them in the same
    order as the face-flux strategy, so the residuals are identical."""
    torch = torch_cuda
    solver = case.solver()
    U = torch.tensor(case.initial_state(), dtype=torch.float64, device="cuda")
    Ra, Rb = torch.zeros_like(U), torch.zeros_like(U)
    solver.set_strategy("face_flux")
    solver.compute_rhs(0.0, U, Ra)
    solver.set_strategy("tile")
    solver.compute_rhs(0.0, U, Rb)
    torch.cuda.synchronize()
    assert torch.equal(Ra, Rb), (Ra - Rb).abs().max().item()


@pytest.mark.parametrize("case", CASES, ids=[c.name for c in CASES])
def test_fused_agrees_with_face_flux(case, torch_cuda):
    """The fused kernel solves each cell's least-squares normal equations instead of applying the stored
    pseudo-inverse weights, so it agrees with the face-flux residual to rounding (not bitwise); it is deterministic
    run to run."""
    torch = torch_cuda
    solver = case.solver()
    use_strategy(solver, "fused")
    U = torch.tensor(case.initial_state(), dtype=torch.float64, device="cuda")
    Ra, Rb, Rc = torch.zeros_like(U), torch.zeros_like(U), torch.zeros_like(U)
    solver.compute_rhs(0.0, U, Ra)
    solver.compute_rhs(0.0, U, Rc)
    solver.set_strategy("face_flux")
    solver.compute_rhs(0.0, U, Rb)
    torch.cuda.synchronize()
    assert torch.equal(Ra, Rc)
    NI = case.mesh.n_interior
    err = ((Ra - Rb)[:, :NI].norm(dim=1) / Rb[:, :NI].norm(dim=1).clamp_min(1e-300)).max().item()
    assert err <= 1e-13, err


MARCH_CASES = CASES + [
    Case("march_columns_periodic", BoxMesh([24, 16, 20], [0, 0, 0], [1, 1, 1], [True] * 3, ordering="column"),
         recon="leastsquares", limiter="barthjespersen", mu=1.8e-5, k=0.0251125),
    Case("march_columns_walls", BoxMesh([20, 12, 16], [0, 0, 0], [1, 1, 1], [False, True, False], ordering="column"),
         recon="leastsquares", limiter="barthjespersen", mu=0.05, k=0.5, bcs={**inflow_x(0), 1: {"kind": "outflow"},
                                                                               4: {"kind": "outflow"}, 5: {"kind": "outflow"}}),
]


@pytest.mark.parametrize("case", MARCH_CASES, ids=[c.name for c in MARCH_CASES])
def test_march_agrees_with_face_flux_and_oracle(case, torch_cuda):
    """The march kernel solves the least-squares normal equations per cell like the fused one, so it agrees with the
    face-flux residual to rounding (not bitwise) and with the oracle to 1e-12; deterministic run to run."""
    torch = torch_cuda
    solver = case.solver()
    use_strategy(solver, "march")
    print(f"[parity] march {case.name}", flush=True)
    U0 = case.initial_state()
    U = torch.tensor(U0, dtype=torch.float64, device="cuda")
    Ra, Rb, Rc = torch.zeros_like(U), torch.zeros_like(U), torch.zeros_like(U)
    solver.compute_rhs(0.0, U, Ra)
    solver.compute_rhs(0.0, U, Rc)
    solver.set_strategy("face_flux")
    solver.compute_rhs(0.0, U, Rb)
    torch.cuda.synchronize()
    assert torch.equal(Ra, Rc)
    NI = case.mesh.n_interior
    err = ((Ra - Rb)[:, :NI].norm(dim=1) / Rb[:, :NI].norm(dim=1).clamp_min(1e-300)).max().item()
    assert err <= 1e-13, err
    Rref, _ = case.oracle_rhs(U0.copy())
    Rg = Ra.cpu().numpy()
    for p in range(Rg.shape[0]):
        ref = Rref[p, :NI]
        e = np.linalg.norm(Rg[p, :NI] - ref) / max(np.linalg.norm(ref), 1e-300)
        assert e <= TOL_FP64, (p, e)


def test_march_available_on_column_boxes(torch_cuda):
    """The march strategy is offered on column-ordered boxes (periodic and walled) and refused with a message for
    species, the mixed form and the PETSc limiter family."""
    for case in MARCH_CASES[-2:]:
        assert case.solver().has_strategy("march"), case.name
    box = MARCH_CASES[-2].mesh
    assert not Case("m_mixed", box, recon="leastsquares", limiter="barthjespersen").solver(form="mixed").has_strategy("march")
    species = Case("m_species", box, recon="leastsquares", limiter="barthjespersen", n_species=2).solver()
    with pytest.raises(RuntimeError, match="march strategy is unavailable"):
        species.set_strategy("march")
    assert not Case("m_minmod", box, recon="leastsquares", limiter="minmod").solver().has_strategy("march")


def test_march_accumulates(torch_cuda):
    """accumulate = 1 adds the residual into R (reference locF += semantics), also for the march kernel."""
    torch = torch_cuda
    case = MARCH_CASES[-2]
    solver = case.solver()
    use_strategy(solver, "march")
    U = torch.tensor(case.initial_state(), dtype=torch.float64, device="cuda")
    R1 = torch.zeros_like(U)
    solver.compute_rhs(0.0, U, R1)
    R2 = torch.full_like(U, 2.5)
    solver.compute_rhs(0.0, U, R2, accumulate=True)
    torch.cuda.synchronize()
    NI = case.mesh.n_interior
    assert torch.equal(R2[:, :NI], R1[:, :NI] + 2.5)


def test_fused_availability(torch_cuda):
    """The fused strategy is offered for FP64 handles without species and refused with a message otherwise."""
    from ablate_amd.mesh import BoxMesh

    box = BoxMesh([8, 8, 8], [0, 0, 0], [1, 1, 1], [False, True, True])
    plain = Case("fused_box", box, recon="leastsquares", limiter="barthjespersen", mu=1.8e-5, k=0.0251125,
                 bcs=inflow_x(0))
    assert plain.solver().has_strategy("fused")
    assert not plain.solver(form="mixed").has_strategy("fused")
    species = Case("fused_species", box, recon="leastsquares", limiter="barthjespersen", n_species=3, bcs=inflow_x(3))
    s = species.solver()
    assert not s.has_strategy("fused")
    with pytest.raises(RuntimeError, match="fused strategy is unavailable"):
        s.set_strategy("fused")
    petsc = Case("fused_minmod", box, recon="leastsquares", limiter="minmod", bcs=inflow_x(0))
    assert not petsc.solver().has_strategy("fused")


@pytest.mark.parametrize("form", ["fp64", "mixed"])
def test_auto_strategy_choice(form, torch_cuda):
    """The default strategy: tiles on a structured box (row tiles evaluate <= ~1.7 faces per face), face flux on a
    tetrahedral mesh in bandwidth-reducing order (~2 evaluations per face) and where species rows fill the LDS; the
    fused strategy only on request."""
    from ablate_amd.mesh import BoxMesh, TetBoxMesh

    box = Case("auto_box", BoxMesh([32, 32, 32], [0, 0, 0], [1, 1, 1], [True] * 3), recon="leastsquares",
               limiter="barthjespersen", mu=1.8e-5, k=0.0251125)
    # 48000 tets in reverse Cuthill-McKee order: a tile of consecutive cells shares almost no faces (1.9 evaluations
    # per face)
    tets = Case("auto_tets", TetBoxMesh(20, 0.2), recon="leastsquares", limiter="barthjespersen", mu=0.05, k=0.5,
                bcs=all_outflow(6))
    species = Case("auto_species", BoxMesh([32, 32, 32], [0, 0, 0], [1, 1, 1], [False, True, True]), n_species=10,
                   bcs=inflow_x(10))
    best = "tile"
    assert box.solver(form=form).strategy() == best
    assert tets.solver(form=form).strategy() == "face_flux"
    assert species.solver(form=form).strategy() == "face_flux"
    s = box.solver(form=form)
    s.set_strategy("face_flux")
    assert s.strategy() == "face_flux"
    s.set_strategy("auto")
    assert s.strategy() == best


@pytest.mark.parametrize("form", ["fp64", "mixed"])
def test_tile_size_does_not_change_the_residual(form, torch_cuda):
    """Tiles of 64 to 512 cells (512: more than 64 KB of LDS per workgroup) give bitwise the face-flux residual."""
    from ablate_amd.mesh import BoxMesh

    torch = torch_cuda
    case = Case("tiles_box", BoxMesh([24, 24, 24], [0, 0, 0], [1, 1, 1], [True] * 3), recon="leastsquares",
                limiter="barthjespersen", mu=1.8e-5, k=0.0251125)
    U = torch.tensor(case.initial_state(), dtype=torch.float64, device="cuda")
    s = case.solver(form=form)
    ref = torch.zeros_like(U)
    s.set_strategy("face_flux")
    s.compute_rhs(0.0, U, ref)
    s.set_strategy("tile")
    for cells in (64, 128, 512, 0):
        print(f"[parity] tile size {cells} {form}", flush=True)
        s.set_tile_size(cells)
        R = torch.zeros_like(U)
        s.compute_rhs(0.0, U, R)
        torch.cuda.synchronize()
        assert torch.equal(R, ref), cells


def test_pack_unpack_roundtrip(torch_cuda):
    torch = torch_cuda
    case = CASES[0]
    solver = case.solver()
    soa = torch.randn(solver.n_comp, case.mesh.n_cells, dtype=torch.float64, device="cuda")
    aos = torch.empty(case.mesh.n_cells, solver.n_comp, dtype=torch.float64, device="cuda")
    import ctypes

    from ablate_amd import _lib

    L = _lib.lib()
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    _lib.check(L.abl_fv_unpack_aos(solver.handle, ctypes.c_void_p(soa.data_ptr()), ctypes.c_void_p(aos.data_ptr()), stream),
               solver.handle, "unpack")
    back = torch.empty_like(soa)
    _lib.check(L.abl_fv_pack_aos(solver.handle, ctypes.c_void_p(aos.data_ptr()), ctypes.c_void_p(back.data_ptr()), stream),
               solver.handle, "pack")
    torch.cuda.synchronize()
    assert torch.equal(aos, soa.t())
    assert torch.equal(back, soa)


def test_check_state_counts_bad_cells(torch_cuda):
    torch = torch_cuda
    case = CASES[1]
    solver = case.solver()
    U = torch.tensor(case.initial_state(), dtype=torch.float64, device="cuda")
    assert solver.check_state(U) == 0
    U[0, 3] = -1.0
    U[2, 7] = float("nan")
    U[1, 9] = float("inf")
    assert solver.check_state(U) == 3
