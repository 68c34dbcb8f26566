"""Hand-written adjoint (abl_fv_rhs_vjp) of the RHS against (1) torch.autograd over independent PyTorch FP64
restatements (tests/torch_rhs.py: first order, and second order with least squares + Barth-Jespersen, viscous and
species diffusion terms; both pinned to the oracle) and (2) central finite differences (h = 1e-6) of the CPU
oracle, including the inflow parameters theta = (rho_in, |u_in|, T_in)."""
import numpy as np
import pytest
import torch as _torch

from ablate_amd import autograd as fv_autograd
from ablate_amd import synthetic
from ablate_amd.mesh import BoxMesh, TetBoxMesh
from tests.fvcases import Case, all_outflow, ghost_and_at_face, inflow_x
from tests.test_torch_restatement import PHASE, _cases2, thetas_of, torch_model, torch_model2

pytestmark = pytest.mark.gpu


def adjoint_cases():
    chan = lambda n: BoxMesh([n, n, n], [0, 0, 0], [1, 1, 1], [False, True, True])
    return [
        Case("c4_ausm_species_inflow", chan(8), n_species=10, bcs=inflow_x(10), phase=PHASE),
        Case("rusanov_species_inflow", chan(6), flux="rusanov", n_species=3, bcs=inflow_x(3), phase=PHASE),
        Case("periodic_ausm", BoxMesh([6, 6, 6], [0, 0, 0], [1, 1, 1], [True] * 3), phase=PHASE),
        Case("quad2d_outflow", BoxMesh([9, 7], [0, 0], [1, 1], [False, True]), state="random2d",
             bcs=all_outflow(2)),
        Case("tets_rusanov_outflow", TetBoxMesh(3, 0.2), flux="rusanov", n_species=2, bcs=all_outflow(6),
             phase=PHASE),
        Case("caller_ghost_at_face_species", chan(6), n_species=3, bcs=ghost_and_at_face(3), phase=PHASE),
        # 4608 boundary faces: the theta_bar reduction's four-loads-in-flight loop (1024-lane blocks) and its tail
        Case("inflow_4608_boundary_faces", BoxMesh([3, 48, 48], [0, 0, 0], [1, 1, 1], [False, True, True]),
             n_species=2, bcs=inflow_x(2), phase=PHASE),
    ]


CASES = adjoint_cases()


def _seed(case, U0):
    return synthetic.adjoint_seed(U0.shape[0], case.mesh.n_cells, case.mesh.n_interior, cells=case.mesh.source_cells())


@pytest.mark.parametrize("case", CASES, ids=lambda c: c.name)
def test_vjp_matches_autograd_of_restatement(case, torch_cuda):
    torch = torch_cuda
    print(f"[adjoint] {case.name}", flush=True)
    U0 = case.initial_state()
    Rb = _seed(case, U0)
    solver = case.solver()
    U = torch.tensor(U0, dtype=torch.float64, device="cuda")
    Ub = torch.empty_like(U)
    theta_bar = solver.rhs_vjp(0.0, U, torch.tensor(Rb, device="cuda"), Ub)
    torch.cuda.synchronize()
    Ug = Ub.cpu().numpy()
    # reference gradient: autograd through the PyTorch restatement
    Ut = _torch.tensor(U0, requires_grad=True)
    thetas = {s: t.clone().requires_grad_(True) for s, t in thetas_of(case).items()}
    out = torch_model(case)(Ut, thetas)
    loss = (out * _torch.tensor(Rb)).sum()
    grads = _torch.autograd.grad(loss, [Ut] + list(thetas.values()))
    ref = grads[0].numpy()
    NI = case.mesh.n_interior
    err = np.linalg.norm(Ug[:, :NI] - ref[:, :NI]) / np.linalg.norm(ref[:, :NI])
    assert err <= 1e-10, err
    # ghost rows: the adjoint of caller-filled ghost values, zero for library-filled ghosts (autograd gives 0 there)
    np.testing.assert_allclose(Ug[:, NI:], ref[:, NI:], rtol=1e-10, atol=1e-12 * np.abs(ref).max())
    for (slot, t), g in zip(thetas.items(), grads[1:]):
        np.testing.assert_allclose(theta_bar[slot], g.numpy(), rtol=1e-10, atol=1e-12 * np.abs(g.numpy()).max())


@pytest.mark.parametrize("case", CASES[:2], ids=lambda c: c.name)
def test_vjp_matches_oracle_central_differences(case, torch_cuda):
    """Rbar . (R(U + hV) - R(U - hV)) / 2h of the oracle vs Ubar . V, and the same for each inflow parameter."""
    torch = torch_cuda
    rng = np.random.default_rng(3)
    U0 = case.initial_state()
    NI = case.mesh.n_interior
    Rb = _seed(case, U0)
    V = np.zeros_like(U0)
    V[:, :NI] = rng.normal(size=(U0.shape[0], NI)) * np.abs(U0[:, :NI]).mean(axis=1, keepdims=True)
    h = 1e-6
    Rp, _ = case.oracle_rhs(U0 + h * V)
    Rm, _ = case.oracle_rhs(U0 - h * V)
    fd = np.sum(Rb * (Rp - Rm)) / (2 * h)
    solver = case.solver()
    U = torch.tensor(U0, dtype=torch.float64, device="cuda")
    Ub = torch.empty_like(U)
    theta_bar = solver.rhs_vjp(0.0, U, torch.tensor(Rb, device="cuda"), Ub)
    ad = float((Ub.cpu().numpy() * V).sum())
    assert ad == pytest.approx(fd, rel=1e-6)
    # inflow parameters (slot 0): rho_in, |u_in|, T_in
    b = case.bcs[0]
    base = np.array([b["rho"], np.linalg.norm(b["vel"]), b["T"]])
    d = np.asarray(b["vel"]) / np.linalg.norm(b["vel"])
    for j in range(3):
        th = base.copy()
        hj = h * abs(base[j])
        vals = []
        for sgn in (+1, -1):
            th[j] = base[j] + sgn * hj
            bcs = dict(case.oracle_bcs())
            bcs[0] = dict(bcs[0], rho=th[0], vel=list(th[1] * d), T=th[2])
            from oracle import binding as orc
            R, _ = orc.rhs(case.mesh.arrays, case.oracle_physics(), U0.copy(), bcs)
            vals.append(np.sum(Rb * R))
        fdj = (vals[0] - vals[1]) / (2 * hj)
        assert theta_bar[0][j] == pytest.approx(fdj, rel=1e-6, abs=1e-9 * abs(fdj) + 1e-12)


def test_autograd_function_wrapper(torch_cuda):
    """torch.autograd through ablate_amd.autograd.rhs returns the hand-written VJP (state and inflow theta)."""
    torch = torch_cuda
    case = CASES[0]
    U0 = case.initial_state()
    Rb = torch.tensor(_seed(case, U0), device="cuda")
    solver = case.solver()
    U = torch.tensor(U0, dtype=torch.float64, device="cuda", requires_grad=True)
    theta = torch.tensor([1.2, 100.0, 300.0], dtype=torch.float64, requires_grad=True)
    R = fv_autograd.rhs(U, solver, theta, slot=0, direction=(1.0, 0.0, 0.0), Yi=case.bcs[0]["Yi"])
    gU, gT = torch.autograd.grad((R * Rb).sum(), [U, theta])
    Ub = torch.empty_like(U)
    tb = solver.rhs_vjp(0.0, U.detach().clone(), Rb, Ub)
    assert torch.equal(gU, Ub)
    np.testing.assert_array_equal(gT.numpy(), tb[0])


def test_vjp_deterministic_and_rejects_unsupported(torch_cuda):
    torch = torch_cuda
    case = CASES[0]
    U0 = case.initial_state()
    solver = case.solver()
    U = torch.tensor(U0, dtype=torch.float64, device="cuda")
    Rb = torch.tensor(_seed(case, U0), device="cuda")
    a, b = torch.empty_like(U), torch.empty_like(U)
    solver.rhs_vjp(0.0, U, Rb, a)
    solver.rhs_vjp(0.0, U, Rb, b)
    torch.cuda.synchronize()
    assert torch.equal(a, b)
    recon = Case("lsq", case.mesh, recon="leastsquares", limiter="vanleer", n_species=10, bcs=case.bcs).solver()
    from ablate_amd._lib import NativeError

    with pytest.raises(NativeError, match="Barth-Jespersen"):
        recon.rhs_vjp(0.0, U, Rb, a)


# ---- second order (least squares + Barth-Jespersen or none, viscous flux, species diffusion)

def adjoint_cases2():
    chan = lambda n: BoxMesh([n, n, n], [0, 0, 0], [1, 1, 1], [False, True, True])
    return _cases2() + [
        Case("c5_like_species10_lsq_bj_rusanov", TetBoxMesh(3, 0.2), flux="rusanov", n_species=10,
             recon="leastsquares", limiter="barthjespersen", mu=1.8e-5, k=0.0251125, diff=2e-5, bcs=all_outflow(6),
             phase=PHASE),
        Case("ns_inflow_ausm_lsq_bj", chan(6), n_species=2, recon="leastsquares", limiter="barthjespersen", mu=0.5,
             k=5.0, diff=0.3, bcs=inflow_x(2), phase=PHASE),
        Case("viscous_first_order", chan(5), mu=0.5, k=5.0, bcs=inflow_x(), phase=PHASE),
    ]


CASES2 = adjoint_cases2()


@pytest.mark.parametrize("case", CASES2, ids=lambda c: c.name)
def test_second_order_vjp_matches_autograd(case, torch_cuda):
    torch = torch_cuda
    print(f"[adjoint2] {case.name}", flush=True)
    U0 = case.initial_state()
    Rb = _seed(case, U0)
    solver = case.solver()
    U = torch.tensor(U0, dtype=torch.float64, device="cuda")
    Ub = torch.empty_like(U)
    theta_bar = solver.rhs_vjp(0.0, U, torch.tensor(Rb, device="cuda"), Ub)
    torch.cuda.synchronize()
    Ug = Ub.cpu().numpy()
    Ut = _torch.tensor(U0, requires_grad=True)
    thetas = {s: t.clone().requires_grad_(True) for s, t in thetas_of(case).items()}
    out = torch_model2(case)(Ut, thetas)
    grads = _torch.autograd.grad((out * _torch.tensor(Rb)).sum(), [Ut] + list(thetas.values()))
    ref = grads[0].numpy()
    NI = case.mesh.n_interior
    err = np.linalg.norm(Ug[:, :NI] - ref[:, :NI]) / np.linalg.norm(ref[:, :NI])
    print(f"[adjoint2] {case.name} rel err vs autograd {err:.3e}", flush=True)
    assert err <= 1e-10, err
    np.testing.assert_allclose(Ug[:, NI:], ref[:, NI:], rtol=1e-9, atol=1e-11 * np.abs(ref).max())
    for (slot, t), g in zip(thetas.items(), grads[1:]):
        np.testing.assert_allclose(theta_bar[slot], g.numpy(), rtol=1e-9, atol=1e-12 * np.abs(g.numpy()).max())


@pytest.mark.parametrize("case", [c for c in CASES2 if c.bcs and 0 in c.bcs and c.bcs[0]["kind"] == "inflow"][:2],
                         ids=lambda c: c.name)
def test_second_order_vjp_matches_oracle_central_differences(case, torch_cuda):
    torch = torch_cuda
    rng = np.random.default_rng(5)
    U0 = case.initial_state()
    NI = case.mesh.n_interior
    Rb = _seed(case, U0)
    V = np.zeros_like(U0)
    V[:, :NI] = rng.normal(size=(U0.shape[0], NI)) * np.abs(U0[:, :NI]).mean(axis=1, keepdims=True)
    h = 1e-6
    Rp, _ = case.oracle_rhs(U0 + h * V)
    Rm, _ = case.oracle_rhs(U0 - h * V)
    fd = np.sum(Rb * (Rp - Rm)) / (2 * h)
    solver = case.solver()
    U = torch.tensor(U0, dtype=torch.float64, device="cuda")
    Ub = torch.empty_like(U)
    solver.rhs_vjp(0.0, U, torch.tensor(Rb, device="cuda"), Ub)
    ad = float((Ub.cpu().numpy() * V).sum())
    assert ad == pytest.approx(fd, rel=1e-5)


def test_second_order_vjp_deterministic(torch_cuda):
    torch = torch_cuda
    case = CASES2[0]
    U0 = case.initial_state()
    solver = case.solver()
    U = torch.tensor(U0, dtype=torch.float64, device="cuda")
    Rb = torch.tensor(_seed(case, U0), device="cuda")
    a, b = torch.empty_like(U), torch.empty_like(U)
    solver.rhs_vjp(0.0, U, Rb, a)
    solver.rhs_vjp(0.0, U, Rb, b)
    torch.cuda.synchronize()
    assert torch.equal(a, b)


def test_second_order_vjp_c5_physics_at_55k_tets(torch_cuda):
    """The C5 adjoint workload's physics (Navier-Stokes + 10 species, mu = 1.8e-5, Pr = 0.72, D = 2e-5, LSQ +
    Barth-Jespersen, Rusanov, outflow ghosts on all sides) on 21^3 x 6 = 55566 jittered tetrahedra: the hand-written
    VJP against torch.autograd of the independent PyTorch restatement to 1e-10."""
    torch = torch_cuda
    mu = 1.8e-5
    case = Case("c5_physics_55k_tets", TetBoxMesh(21, 0.2), flux="rusanov", n_species=10, recon="leastsquares",
                limiter="barthjespersen", mu=mu, k=mu * 1.4 * 287.0 / 0.4 / 0.72, diff=2e-5, bcs=all_outflow(6),
                phase=PHASE)
    print(f"[adjoint2] {case.name} cells={case.mesh.n_interior}", flush=True)
    U0 = case.initial_state()
    Rb = _seed(case, U0)
    solver = case.solver()
    U = torch.tensor(U0, dtype=torch.float64, device="cuda")
    Ub = torch.empty_like(U)
    solver.rhs_vjp(0.0, U, torch.tensor(Rb, device="cuda"), Ub)
    torch.cuda.synchronize()
    Ug = Ub.cpu().numpy()
    Ut = _torch.tensor(U0, requires_grad=True)
    out = torch_model2(case)(Ut, {})
    (ref,) = _torch.autograd.grad((out * _torch.tensor(Rb)).sum(), [Ut])
    ref = ref.numpy()
    NI = case.mesh.n_interior
    err = np.linalg.norm(Ug[:, :NI] - ref[:, :NI]) / np.linalg.norm(ref[:, :NI])
    print(f"[adjoint2] {case.name} rel err vs autograd {err:.3e}", flush=True)
    assert err <= 1e-10, err
    # per field as well: the species rows are small next to the euler rows
    for p in range(U0.shape[0]):
        ep = np.linalg.norm(Ug[p, :NI] - ref[p, :NI]) / np.linalg.norm(ref[p, :NI])
        assert ep <= 1e-10, (p, ep)


def test_vjp_off_flux_is_zero_everywhere(torch_cuda):
    """OffFlux without diffusion terms: R = 0 for every state, so Ubar = 0 in every row (ghost rows included), written
    by a fill kernel over a NaN-filled output."""
    torch = torch_cuda
    chan = BoxMesh([6, 5, 4], [0, 0, 0], [1, 1, 1], [False, True, True])
    case = Case("off_species_inflow", chan, flux="off", n_species=2, bcs=inflow_x(2), phase=PHASE)
    U0 = case.initial_state()
    solver = case.solver()
    U = torch.tensor(U0, dtype=torch.float64, device="cuda")
    Ub = torch.full_like(U, float("nan"))
    theta_bar = solver.rhs_vjp(0.0, U, torch.tensor(_seed(case, U0), device="cuda"), Ub)
    torch.cuda.synchronize()
    assert torch.count_nonzero(Ub).item() == 0 and not torch.isnan(Ub).any().item()
    assert not np.any(np.asarray(theta_bar))
