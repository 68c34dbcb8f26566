"""The CPU oracle against the reference's own unit-test tables (tests/golden/*.json, extracted from the ABLATE
tests by tests/golden/make_golden.py) and against what the mathematics fixes."""
import json
import math
from pathlib import Path

import numpy as np
import pytest

from oracle import binding as orc

GOLDEN = Path(__file__).resolve().parent / "golden"
DIRS = {"LEFT": orc.LEFT, "RIGHT": orc.RIGHT, "NA": orc.NA}


def _load(name):
    return json.loads((GOLDEN / name).read_text())


@pytest.mark.parametrize("name,fn", [("Ausm", orc.ausm), ("AverageFlux", orc.average_flux), ("OffFlux", orc.off_flux)])
def test_flux_calculator_tables(name, fn):
    """fluxCalculatorTests.cpp: direction exact, mass flux within 1e-2, interface pressure within 1."""
    t = _load("flux_calculators.json")[name]
    n = len(t["uL"])
    assert n == len(t["expectedDirection"])
    for i in range(n):
        d, m, p = fn(t["uL"][i], t["aL"][i], t["rhoL"][i], t["pL"][i], t["uR"][i], t["aR"][i], t["rhoR"][i], t["pR"][i])
        assert d == DIRS[t["expectedDirection"][i]], (name, i)
        assert abs(m - t["expectedMassFlux"][i]) <= t["tolerance"]["massFlux"], (name, i, m)
        assert abs(p - t["expectedInterfacePressure"][i]) <= t["tolerance"]["interfacePressure"], (name, i, p)


def test_ausm_closed_form_and_invariants():
    """Liou-Steffen polynomials: M+(M) + M-(M) = M and P+(M) + P-(M) = 1 for |M| <= 1; supersonic upwinding; the
    split is antisymmetric under swapping the sides with the normal reversed."""
    rng = np.random.default_rng(3)
    for _ in range(200):
        a = rng.uniform(50, 400)
        M = rng.uniform(-0.99, 0.99)
        rho, p = rng.uniform(0.5, 3), rng.uniform(1e4, 1e6)
        # same state both sides: mdot = M a rho (mass flux of the state), p12 = p
        d, m, p12 = orc.ausm(M * a, a, rho, p, M * a, a, rho, p)
        assert m == pytest.approx(M * a * rho, rel=1e-12, abs=1e-9)
        assert p12 == pytest.approx(p, rel=1e-12)
        # supersonic: full upwinding
        Ms = rng.uniform(1.01, 3.0)
        d, m, p12 = orc.ausm(Ms * a, a, rho, p, Ms * a * 1.1, a, rho * 2, p * 3)
        assert d == orc.LEFT and m == pytest.approx(Ms * a * rho, rel=1e-12) and p12 == pytest.approx(p, rel=1e-12)
        # antisymmetry: F(L, R, n) = -F(R, L, -n)
        uL, uR = rng.uniform(-300, 300, 2)
        aL, aR = rng.uniform(100, 400, 2)
        rL, rR = rng.uniform(0.5, 2, 2)
        pL, pR = rng.uniform(1e4, 1e6, 2)
        d1, m1, q1 = orc.ausm(uL, aL, rL, pL, uR, aR, rR, pR)
        d2, m2, q2 = orc.ausm(-uR, aR, rR, pR, -uL, aL, rL, pL)
        assert m1 == pytest.approx(-m2, rel=1e-13, abs=1e-12)
        assert q1 == pytest.approx(q2, rel=1e-13)


def test_ausm_planted_mistake_fails_table():
    """A plausible slip (sign of the subsonic M- branch) would break the reference table: checks that the table
    pins the branch."""
    t = _load("flux_calculators.json")["Ausm"]

    def wrong(uL, aL, rhoL, pL, uR, aR, rhoR, pR):
        Mm = uR / aR
        sMm = 0.25 * (Mm - 1) ** 2 if abs(Mm) <= 1 else 0.5 * (Mm - abs(Mm))  # planted: sign flipped
        Mp = uL / aL
        sMp = 0.25 * (Mp + 1) ** 2 if abs(Mp) <= 1 else 0.5 * (Mp + abs(Mp))
        m = sMm + sMp
        return m * (aR * rhoR if m < 0 else aL * rhoL)

    bad = sum(abs(wrong(*[t[k][i] for k in ("uL", "aL", "rhoL", "pL", "uR", "aR", "rhoR", "pR")]) - t["expectedMassFlux"][i]) > 1e-2
              for i in range(len(t["uL"])))
    assert bad > 10


def test_ns_advection_flux_table():
    """navierStokesTransportTests.cpp:56-81 (1-D AdvectionFlux with AUSM, PerfectGas defaults, tolerance 1e-3)."""
    for case in _load("ns_advection_flux.json"):
        out = orc.ns_advection_flux(case["area"], case["xLeft"], case["xRight"], case["eos"]["gamma"],
                                    case["eos"]["Rgas"], case["fluxCalculator"].lower())
        np.testing.assert_allclose(out, case["expectedFlux"], atol=case["tolerance"], rtol=0)


def test_ns_advection_flux_symmetry_3d():
    """Swapping the sides with the normal reversed negates the flux; a uniform state gives F.n of the state."""
    rng = np.random.default_rng(5)
    g, R = 1.4, 287.0
    for _ in range(50):
        n = rng.normal(size=3) * rng.uniform(0.1, 2)
        xl = np.array([rng.uniform(0.5, 2), 0, *rng.uniform(-100, 100, 3)])
        xr = np.array([rng.uniform(0.5, 2), 0, *rng.uniform(-100, 100, 3)])
        for x in (xl, xr):
            p = rng.uniform(5e4, 2e5)
            x[1] = p / (g - 1) + 0.5 * np.dot(x[2:], x[2:]) / x[0]
        f1 = orc.ns_advection_flux(n, xl, xr)
        f2 = orc.ns_advection_flux(-n, xr, xl)
        np.testing.assert_allclose(f1, -f2, rtol=1e-12, atol=1e-9 * np.abs(f1).max())
        # uniform state: physical flux
        f = orc.ns_advection_flux(n, xl, xl)
        rho, u = xl[0], xl[2:] / xl[0]
        p = (g - 1) * (xl[1] - 0.5 * rho * np.dot(u, u))
        un = np.dot(u, n)
        exact = np.concatenate([[rho * un, (xl[1] + p) * un], rho * u * un + p * n])
        np.testing.assert_allclose(f, exact, rtol=1e-11, atol=1e-9 * np.abs(exact).max())


def test_stress_tensor_table():
    for case in _load("stress_tensor.json"):
        out = orc.stress_tensor(case["dim"], case["mu"], case["gradVel"])
        np.testing.assert_allclose(out, case["expectedStressTensor"], atol=case["tolerance"], rtol=0)


_PG_PROPS = {
    "Pressure": "pressure", "InternalSensibleEnergy": "internal_energy", "SpeedOfSound": "speed_of_sound",
    "Temperature": "temperature", "SpecificHeatConstantPressure": "cp", "SpecificHeatConstantVolume": "cv",
    "Density": "density", "SensibleEnthalpy": "sensible_enthalpy",
}


def test_perfect_gas_table():
    """perfectGasTests.cpp:91-236: each property directly and through the computed temperature."""
    for case in _load("perfect_gas.json"):
        off, nco = case["eulerOffset"], case["eulerComponents"]
        euler = case["conservedValues"][off : off + nco]
        g, R = case["gamma"], case["Rgas"]
        T = orc.pg("temperature", euler, g, R)
        if case["temperature"] is not None:
            assert T == pytest.approx(case["temperature"], abs=1e-6)
        prop = case["property"]
        if prop == "SpeciesSensibleEnthalpy":
            assert case["expectedValue"] == [0.0] * len(case["species"])  # PerfectGas h_i = 0
            continue
        name = _PG_PROPS[prop]
        direct = orc.pg(name, euler, g, R)
        assert direct == pytest.approx(case["expectedValue"][0], abs=case["tolerance"]), case["source"]
        if name in ("internal_energy", "speed_of_sound", "sensible_enthalpy"):
            viaT = orc.pg(name, euler, g, R, T=T)
            assert viaT == pytest.approx(case["expectedValue"][0], abs=case["tolerance"]), case["source"]


def test_lsq_weights_match_pseudoinverse():
    """PetscFVComputeGradient returns the least-squares pseudo-inverse: compare with LAPACK's SVD pinv, including
    rank-deficient stencils."""
    rng = np.random.default_rng(11)
    for dim in (2, 3):
        for n in range(1, 9):
            dx = rng.normal(size=(n, dim))
            w = orc.lsq_weights(dx)
            np.testing.assert_allclose(w, np.linalg.pinv(dx).T, rtol=1e-9, atol=1e-12)
    # rank deficient: colinear rows
    dx = np.outer(rng.normal(size=4), [1.0, 2.0, -1.0])
    np.testing.assert_allclose(orc.lsq_weights(dx), np.linalg.pinv(dx).T, rtol=1e-9, atol=1e-12)
    # uniform hex closed form: w = dx / (2 h^2) per axis
    h = 0.1
    dx = np.array([[h, 0, 0], [-h, 0, 0], [0, h, 0], [0, -h, 0], [0, 0, h], [0, 0, -h]])
    np.testing.assert_allclose(orc.lsq_weights(dx), dx / (2 * h * h), rtol=1e-13)


@pytest.mark.parametrize("name", ["minmod", "vanleer", "vanalbada", "superbee", "mc", "sin"])
def test_limiters_symmetric_form(name):
    """Symmetric-form limiters: phi(1/2) = 1 (linear data unlimited), phi(f <= 0) = 0 (extrema clipped),
    phi(f) = phi(1 - f) on [0, 1], phi in [0, 2]."""
    assert orc.limiter(name, 0.5) == pytest.approx(1.0, abs=1e-15)
    for f in (-3.0, -0.5, -1e-9, 0.0):
        assert orc.limiter(name, f) == pytest.approx(0.0, abs=1e-15)
    for f in np.linspace(0, 1, 41):
        assert orc.limiter(name, f) == pytest.approx(orc.limiter(name, 1 - f), abs=1e-12)
        assert -1e-15 <= orc.limiter(name, f) <= 2.0
    assert orc.limiter("none", 0.3) == 1.0 and orc.limiter("zero", 0.3) == 0.0
