"""The PyTorch restatement used as the independent reference gradient reproduces the oracle's forward RHS, and its
autograd gradient matches central finite differences of the oracle (CPU only)."""
import numpy as np
import pytest
import torch

from ablate_amd.mesh import BoxMesh, TetBoxMesh
from tests.fvcases import Case, all_outflow, ghost_and_at_face, inflow_x, rel_err_per_field
from tests.torch_rhs import TorchRHS


PHASE = (0.37, 0.21, 0.53)  # keeps cell centres off the Taylor-Green symmetry planes (no AUSM upwind ties)


def _cases():
    chan = BoxMesh([6, 5, 4], [0, 0, 0], [1, 1, 1], [False, True, True])
    return [
        Case("c4_like", chan, n_species=3, bcs=inflow_x(3), phase=PHASE),
        Case("c4_like_rusanov", chan, flux="rusanov", n_species=2, bcs=inflow_x(2), phase=PHASE),
        Case("periodic_ausm", BoxMesh([5, 5, 5], [0, 0, 0], [1, 1, 1], [True] * 3), phase=PHASE),
        Case("tets_rusanov", TetBoxMesh(3, 0.2), flux="rusanov", n_species=2, bcs=all_outflow(6), phase=PHASE),
        Case("caller_ghost_at_face", chan, n_species=2, bcs=ghost_and_at_face(2), phase=PHASE),
    ]


def thetas_of(case):
    out = {}
    for slot, b in case.bcs.items():
        if b["kind"] == "inflow":
            um = float(np.linalg.norm(b["vel"]))
            out[slot] = torch.tensor([b["rho"], um, b["T"]], dtype=torch.float64)
    return out


def torch_bcs(case):
    """Boundary slots of a case in the restatements' form (inflow direction / Yi / at-face, outflow, caller ghost)."""
    bcs = {}
    for slot, b in case.bcs.items():
        if b["kind"] == "inflow":
            um = float(np.linalg.norm(b["vel"]))
            bcs[slot] = {"kind": "inflow", "dir": [v / um for v in b["vel"]], "Yi": b.get("Yi"),
                         "at_face": bool(b.get("enforce_at_face", False))}
        elif b["kind"] == "ghost":
            bcs[slot] = {"kind": "ghost"}
        else:
            bcs[slot] = {"kind": "outflow"}
    return bcs


def torch_model(case):
    return TorchRHS(case.mesh, case.gamma, case.Rgas, case.n_species, case.flux, torch_bcs(case))


@pytest.mark.parametrize("case", _cases(), ids=lambda c: c.name)
def test_torch_forward_matches_oracle(case):
    U0 = case.initial_state()
    Rref, _ = case.oracle_rhs(U0.copy())
    Rt = torch_model(case)(torch.tensor(U0), thetas_of(case)).numpy()
    assert rel_err_per_field(Rt, Rref, case.mesh.n_interior).max() <= 1e-12


@pytest.mark.parametrize("case", _cases(), ids=lambda c: c.name)
def test_torch_gradient_matches_oracle_fd(case):
    """Directional derivative of Rbar . R along a random V: autograd vs central differences (h = 1e-6) of the
    oracle."""
    rng = np.random.default_rng(7)
    U0 = case.initial_state()
    NI = case.mesh.n_interior
    V = np.zeros_like(U0)
    V[:, :NI] = rng.normal(size=(U0.shape[0], NI)) * np.abs(U0[:, :NI]).mean(axis=1, keepdims=True)
    Rbar = np.zeros_like(U0)
    Rbar[:, :NI] = rng.normal(size=(U0.shape[0], NI))
    h = 1e-6
    Rp, _ = case.oracle_rhs(U0 + h * V)
    Rm, _ = case.oracle_rhs(U0 - h * V)
    fd = np.sum(Rbar * (Rp - Rm)) / (2 * h)
    U = torch.tensor(U0, requires_grad=True)
    out = torch_model(case)(U, thetas_of(case))
    (g,) = torch.autograd.grad((out * torch.tensor(Rbar)).sum(), U)
    ad = float((g.numpy() * V).sum())
    assert ad == pytest.approx(fd, rel=1e-6)


# ---- second order: least squares + Barth-Jespersen, viscous flux, species advection and diffusion

def _cases2():
    from tests.torch_rhs import TorchRHS2  # noqa: F401  (import check)

    chan = BoxMesh([6, 5, 4], [0, 0, 0], [1, 1, 1], [False, True, True])
    return [
        Case("ns_species_lsq_bj_inflow", chan, n_species=3, recon="leastsquares", limiter="barthjespersen", mu=0.5,
             k=5.0, diff=0.3, bcs=inflow_x(3), phase=PHASE),
        Case("ns_periodic_lsq_bj", BoxMesh([5, 5, 5], [0, 0, 0], [1, 1, 1], [True] * 3), recon="leastsquares",
             limiter="barthjespersen", mu=0.3, k=2.0, phase=PHASE),
        Case("c5_like_tets_rusanov", TetBoxMesh(3, 0.2), flux="rusanov", n_species=2, recon="leastsquares",
             limiter="barthjespersen", mu=0.05, k=0.5, diff=2e-2, bcs=all_outflow(6), phase=PHASE),
        Case("walls_unlimited", BoxMesh([4, 5, 3], [0, 0, 0], [1, 2, 3], [False] * 3), recon="leastsquares",
             limiter="none", mu=0.3, k=2.0, bcs=all_outflow(6), phase=PHASE),
        Case("quad2d_species", BoxMesh([7, 6], [0, 0], [2, 1], [False, True]), n_species=2, recon="leastsquares",
             limiter="barthjespersen", mu=0.2, k=1.0, diff=0.1, bcs=inflow_x(2), phase=PHASE),
        Case("viscous_first_order", chan, n_species=2, mu=0.5, k=5.0, diff=0.3, bcs=inflow_x(2), phase=PHASE),
        Case("caller_ghost_at_face_lsq_bj", chan, n_species=2, recon="leastsquares", limiter="barthjespersen", mu=0.5,
             k=5.0, diff=0.3, bcs=ghost_and_at_face(2), phase=PHASE),
    ]


def torch_model2(case):
    from tests.torch_rhs import TorchRHS2

    return TorchRHS2(case.mesh, case.gamma, case.Rgas, case.n_species, case.flux, torch_bcs(case), case.limiter,
                     case.mu, case.k, case.diff, recon=case.recon == "leastsquares")


@pytest.mark.parametrize("case", _cases2(), ids=lambda c: c.name)
def test_torch_second_order_forward_matches_oracle(case):
    U0 = case.initial_state()
    Rref, _ = case.oracle_rhs(U0.copy())
    Rt = torch_model2(case)(torch.tensor(U0), thetas_of(case)).numpy()
    assert rel_err_per_field(Rt, Rref, case.mesh.n_interior).max() <= 1e-12


@pytest.mark.parametrize("case", _cases2(), ids=lambda c: c.name)
def test_torch_second_order_gradient_matches_oracle_fd(case):
    rng = np.random.default_rng(11)
    U0 = case.initial_state()
    NI = case.mesh.n_interior
    V = np.zeros_like(U0)
    V[:, :NI] = rng.normal(size=(U0.shape[0], NI)) * np.abs(U0[:, :NI]).mean(axis=1, keepdims=True)
    Rbar = np.zeros_like(U0)
    Rbar[:, :NI] = rng.normal(size=(U0.shape[0], NI))
    h = 1e-6
    Rp, _ = case.oracle_rhs(U0 + h * V)
    Rm, _ = case.oracle_rhs(U0 - h * V)
    fd = np.sum(Rbar * (Rp - Rm)) / (2 * h)
    U = torch.tensor(U0, requires_grad=True)
    out = torch_model2(case)(U, thetas_of(case))
    (g,) = torch.autograd.grad((out * torch.tensor(Rbar)).sum(), U)
    ad = float((g.numpy() * V).sum())
    assert ad == pytest.approx(fd, rel=1e-5)
