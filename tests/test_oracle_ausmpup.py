"""The oracle's AUSM+up flux calculator and pressure-gradient-scaling update, pinned by the reference's own tables
(fluxCalculatorTests.cpp AusmpUp rows, tests/golden/ausmpup.json) and by closed forms of the published scheme
(Liou 2006) for the branches those tables leave open (mInf > 0, the PGS alpha update)."""
import json
import math
from pathlib import Path

import numpy as np
import pytest

from oracle import binding as orc

GOLDEN = Path(__file__).resolve().parent / "golden"
DIRS = {"LEFT": orc.LEFT, "RIGHT": orc.RIGHT, "NA": orc.NA}
KEYS = ("uL", "aL", "rhoL", "pL", "uR", "aR", "rhoR", "pR")
TABLES = json.loads((GOLDEN / "ausmpup.json").read_text())


@pytest.mark.parametrize("name", sorted(TABLES))
def test_ausmpup_reference_tables(name):
    """Direction exact; mass flux and interface pressure within the reference's tolerances (1e-2, 1) and, since
    the tables print the reference's own doubles, within 1e-14 relative."""
    t = TABLES[name]
    alpha = t["pgsAlpha"] or 1.0
    for i in range(len(t["uL"])):
        d, m, p = orc.ausmpup(*[t[k][i] for k in KEYS], mInf=t["mInf"], pgs_alpha=alpha)
        em, ep = t["expectedMassFlux"][i], t["expectedInterfacePressure"][i]
        assert d == DIRS[t["expectedDirection"][i]], (name, i)
        assert abs(m - em) <= t["tolerance"]["massFlux"] and abs(p - ep) <= t["tolerance"]["interfacePressure"]
        assert m == pytest.approx(em, rel=1e-14) and p == pytest.approx(ep, rel=1e-14), (name, i, m, p)


def test_ausmpup_uniform_state():
    """Same state on both sides: the split polynomials sum to the state's Mach number (M4+ + M4- = M) and to one
    (P5+ + P5- = 1) for any fa, so mdot = rho u and p12 = p / alpha^2, at every mInf and alpha."""
    rng = np.random.default_rng(5)
    for _ in range(300):
        a, rho, p = rng.uniform(200, 400), rng.uniform(0.5, 3), rng.uniform(1e4, 1e6)
        u = rng.uniform(-1.5, 1.5) * a
        minf = rng.choice([0.0, 0.05, 0.3, 1.0])
        alpha = rng.choice([1.0, 2.0, 30.0])
        d, m, p12 = orc.ausmpup(u, a, rho, p, u, a, rho, p, mInf=minf, pgs_alpha=alpha)
        assert m == pytest.approx(rho * u, rel=1e-12, abs=1e-9 * rho * a)
        assert p12 == pytest.approx(p / alpha**2, rel=1e-12)
        assert d == (orc.LEFT if m > 0 else orc.RIGHT)


def test_ausmpup_pressure_diffusion_closed_form():
    """Fluid at rest with a pressure jump: only the pressure-diffusion term drives the mass flux,
    M_1/2 = -(Kp / fa) (pR - pL) / (rho_1/2 a_1/2^2 alpha^2), with M_o = mInf (M_bar = 0) and fa = M_o (2 - M_o);
    p_1/2 = (pL + pR) / 2 / alpha^2 (P5+-(0) = 1/2).  Pins the mInf branch, the sign of the term and its upwinding."""
    for minf in (0.0, 0.05, 0.3, 0.9):
        fa = 1.0 if minf == 0.0 else minf * (2 - minf)
        for alpha in (1.0, 4.0):
            aL, aR, rL, rR, pL, pR = 340.0, 350.0, 1.1, 1.3, 1.02e5, 0.98e5
            a12 = 0.5 * (aL + aR) / alpha
            r12 = 0.5 * (rL + rR)
            M = -(0.25 / fa) * (pR - pL) / (r12 * a12 * a12 * alpha * alpha)
            d, m, p12 = orc.ausmpup(0.0, aL, rL, pL, 0.0, aR, rR, pR, mInf=minf, pgs_alpha=alpha)
            assert d == orc.LEFT  # pL > pR pushes fluid to the right
            assert m == pytest.approx(a12 * M * rL, rel=1e-13)
            assert p12 == pytest.approx(0.5 * (pL + pR) / alpha**2, rel=1e-13)
            d, m, _ = orc.ausmpup(0.0, aL, rL, pR, 0.0, aR, rR, pL, mInf=minf, pgs_alpha=alpha)
            assert d == orc.RIGHT and m == pytest.approx(-a12 * M * rR, rel=1e-13)


def test_ausmpup_minf_irrelevant_when_fast():
    """fa = 1 whenever max(M_bar^2, mInf^2) >= 1: mInf = 0, mInf = 1 and any mInf at M_bar >= 1 agree exactly."""
    rng = np.random.default_rng(9)
    for _ in range(100):
        a = rng.uniform(200, 400)
        args = (rng.uniform(1.1, 3) * a, a, 1.2, 1e5, rng.uniform(1.1, 3) * a, a * 1.1, 1.4, 1.3e5)
        ref = orc.ausmpup(*args, mInf=0.0)
        assert orc.ausmpup(*args, mInf=0.4) == ref
        slow = (0.1 * a, a, 1.2, 1e5, 0.12 * a, a, 1.3, 1.1e5)
        assert orc.ausmpup(*slow, mInf=1.0) == orc.ausmpup(*slow, mInf=0.0)


# ------------------------------------------------------------------------------ PressureGradientScaling update

def _state(mesh, rho, u, p, gamma=1.4):
    """SoA euler state with per-cell density, x-velocity and pressure (arrays over interior cells)."""
    a = mesh.arrays
    N, NI, dim = a.n_cells, a.n_interior, a.dim
    U = np.zeros((dim + 2, N))
    U[0, :NI] = rho
    U[2, :NI] = rho * u
    U[1, :NI] = p / (gamma - 1) + 0.5 * rho * u * u
    U[:, NI:] = U[:, :1]
    return U


@pytest.fixture(scope="module")
def box():
    from ablate_amd.mesh import BoxMesh

    return BoxMesh([4, 4, 4], [0, 0, 0], [1, 1, 1], [True] * 3)


def test_pgs_update_caps(box):
    """Uniform pressure (max dp = 0): the raw update is huge, so alpha = max(1, min(1.1 alpha_old, 0.7 / Mach_max,
    maxAlphaAllowed)) (pressureGradientScaling.cpp:126-130)."""
    ph = orc.physics()
    NI = box.n_interior
    a0 = math.sqrt(1.4 * 1e5 / 1.2)
    U = _state(box, np.full(NI, 1.2), np.full(NI, 0.01 * a0), np.full(NI, 1e5))
    for alpha_old, max_allowed, want in ((2.0, 100.0, 2.2), (100.0, 100.0, 70.0), (100.0, 50.0, 50.0),
                                         (0.5, 100.0, 1.0)):
        alpha, mach = orc.pgs_update(box.arrays, ph, U, dt=1e-3, alpha=alpha_old, domain_length=1.0,
                                     max_alpha_allowed=max_allowed)
        assert mach == pytest.approx(0.01, rel=1e-12)
        assert alpha == pytest.approx(want, rel=1e-12), (alpha_old, max_allowed)


def test_pgs_update_relaxation_closed_form(box):
    """Fluid at rest with two pressure levels: pRef = mean, max dp = |p2 - p1| / 2, c_max from the higher pressure,
    no Mach cap; alpha = alpha_old + dt c_max / (2 L) (sqrt(fac pRef / max dp) - 1)."""
    ph = orc.physics()
    NI = box.n_interior
    p = np.where(np.arange(NI) % 2 == 0, 1.0e5, 1.2e5)
    U = _state(box, np.full(NI, 1.2), np.zeros(NI), p)
    cmax = math.sqrt(1.4 * 1.2e5 / 1.2)
    pref, dp = 1.1e5, 0.1e5
    for fac, L, dt in ((0.05, 1.0, 1e-4), (0.5, 0.2, 2e-5)):
        alpha, mach = orc.pgs_update(box.arrays, ph, U, dt=dt, alpha=3.0, domain_length=L, max_delta_pressure_fac=fac)
        want = 3.0 + 0.5 * dt * cmax / L * (math.sqrt(fac * pref / (dp + 1e-30)) - 1.0)
        assert mach == 0.0
        assert alpha == pytest.approx(want, rel=1e-13)


def test_pgs_update_nan_domain_length(box):
    """domainLength NaN (the reference's unit-test construction): the NaN update falls through PetscMin's comparison
    to the 10 % growth cap."""
    ph = orc.physics()
    NI = box.n_interior
    U = _state(box, np.full(NI, 1.2), np.full(NI, 1.0), np.full(NI, 1e5))
    alpha, _ = orc.pgs_update(box.arrays, ph, U, dt=1e-3, alpha=5.0, domain_length=float("nan"))
    assert alpha == pytest.approx(5.5, rel=1e-15)
