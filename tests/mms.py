"""Method of manufactured solutions for the compressible RHS, as the reference's residual-convergence test runs it
(tests/unitTests/finiteVolume/compressibleFlowMmsSourceTests.cpp:533-650, rate table :656-895).

The exact solution is the reference's trigonometric field (EulerExact, :69-140): each primitive (rho, u, v, w, p) is
phiO + a sum of sin / cos terms with wave numbers aPhi * pi / L.  The source term S = div F_inviscid - div F_viscous
that makes it a steady solution is derived here from the mathematics, not transcribed: every primitive is evaluated
as a second-order jet (value, gradient, Hessian), the fluxes and their divergences follow from the jet product and
quotient rules, and ``source_fd`` cross-checks the result by fourth-order central differences of the flux field.

Test infrastructure only.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


# ------------------------------------------------------------------------------------------- second-order jets
class Jet:
    """f(x) with its gradient [n, 3] and Hessian [n, 3, 3] at n points."""

    __slots__ = ("v", "g", "h")

    def __init__(self, v, g, h):
        self.v, self.g, self.h = v, g, h

    @staticmethod
    def const(c, n):
        return Jet(np.full(n, float(c)), np.zeros((n, 3)), np.zeros((n, 3, 3)))

    def __add__(self, o):
        if not isinstance(o, Jet):
            return Jet(self.v + o, self.g, self.h)
        return Jet(self.v + o.v, self.g + o.g, self.h + o.h)

    __radd__ = __add__

    def __neg__(self):
        return Jet(-self.v, -self.g, -self.h)

    def __sub__(self, o):
        return self + (-o)

    def __rsub__(self, o):
        return (-self) + o

    def __mul__(self, o):
        if not isinstance(o, Jet):
            return Jet(self.v * o, self.g * o, self.h * o)
        gg = np.einsum("ni,nj->nij", self.g, o.g)
        return Jet(self.v * o.v, self.g * o.v[:, None] + o.g * self.v[:, None],
                   self.h * o.v[:, None, None] + o.h * self.v[:, None, None] + gg + gg.transpose(0, 2, 1))

    __rmul__ = __mul__

    def recip(self):
        iv = 1.0 / self.v
        g = -self.g * (iv * iv)[:, None]
        h = -self.h * (iv * iv)[:, None, None] + 2.0 * np.einsum("ni,nj->nij", self.g, self.g) * (iv ** 3)[:, None, None]
        return Jet(iv, g, h)

    def __truediv__(self, o):
        if not isinstance(o, Jet):
            return self * (1.0 / o)
        return self * o.recip()


def _trig(fn, amp, k, x, axis):
    """amp * fn(k x[axis]) as a jet (fn = sin or cos)."""
    n = x.shape[0]
    t = k * x[:, axis]
    s, c = np.sin(t), np.cos(t)
    v = amp * (s if fn == "sin" else c)
    d1 = amp * k * (c if fn == "sin" else -s)
    d2 = -amp * k * k * (s if fn == "sin" else c)
    g = np.zeros((n, 3))
    h = np.zeros((n, 3, 3))
    g[:, axis] = d1
    h[:, axis, axis] = d2
    return Jet(v, g, h)


# ------------------------------------------------------------------------------------------- the exact field
@dataclass
class Phi:
    """phiO, phiX, phiY, phiZ, aPhiX, aPhiY, aPhiZ of one primitive (the reference's PhiConstants)."""

    o: float = 0.0
    x: float = 0.0
    y: float = 0.0
    z: float = 0.0
    ax: float = 0.0
    ay: float = 0.0
    az: float = 0.0


# which trigonometric function multiplies each direction's term, per primitive (EulerExact, :120-137)
_FORMS = {
    "rho": ("sin", "cos", "sin"),
    "u": ("sin", "cos", "cos"),
    "v": ("cos", "sin", "sin"),
    "w": ("sin", "sin", "cos"),
    "p": ("cos", "sin", "cos"),
}


@dataclass
class MmsConstants:
    dim: int
    rho: Phi
    u: Phi
    v: Phi
    w: Phi = field(default_factory=Phi)
    p: Phi = field(default_factory=Phi)
    L: float = 1.0
    gamma: float = 1.4
    R: float = 287.0
    mu: float = 0.0
    k: float = 0.0


def primitive(c: MmsConstants, name: str, x: np.ndarray) -> Jet:
    ph = getattr(c, name)
    fx, fy, fz = _FORMS[name]
    kk = np.pi / c.L
    out = Jet.const(ph.o, x.shape[0])
    out = out + _trig(fx, ph.x, ph.ax * kk, x, 0)
    out = out + _trig(fy, ph.y, ph.ay * kk, x, 1)
    if c.dim > 2:
        out = out + _trig(fz, ph.z, ph.az * kk, x, 2)
    return out


def _prims(c: MmsConstants, x: np.ndarray):
    x3 = np.zeros((x.shape[0], 3))
    x3[:, : x.shape[1]] = x[:, :3] if x.shape[1] >= 3 else x
    vel = [primitive(c, n, x3) for n in ("u", "v", "w")[: c.dim]]
    return primitive(c, "rho", x3), vel, primitive(c, "p", x3)


def exact_conserved(c: MmsConstants, x: np.ndarray) -> np.ndarray:
    """Conserved state [dim + 2, n] (rho, rhoE, rhoU..) of the exact field at points x [n, >= dim]."""
    rho, vel, p = _prims(c, x)
    ke = sum((u.v * u.v for u in vel), np.zeros_like(rho.v))
    out = np.empty((c.dim + 2, rho.v.size))
    out[0] = rho.v
    out[1] = rho.v * (p.v / ((c.gamma - 1.0) * rho.v) + 0.5 * ke)
    for d in range(c.dim):
        out[2 + d] = rho.v * vel[d].v
    return out


def _flux_parts(c: MmsConstants, x: np.ndarray):
    """Jets of the physical flux components F[comp][d] (inviscid minus viscous) at x."""
    dim = c.dim
    rho, vel, p = _prims(c, x)
    ke = vel[0] * vel[0]
    for d in range(1, dim):
        ke = ke + vel[d] * vel[d]
    rhoE = p / (c.gamma - 1.0) + rho * ke * 0.5
    T = p / (rho * c.R)
    F = [[None] * dim for _ in range(dim + 2)]
    for d in range(dim):
        m = rho * vel[d]
        F[0][d] = m
        F[1][d] = (rhoE + p) * vel[d]
        for cc in range(dim):
            F[2 + cc][d] = m * vel[cc] + (p if cc == d else 0.0)
    return F, vel, T


def source(c: MmsConstants, x: np.ndarray) -> np.ndarray:
    """S = div F_inviscid - div F_viscous of the exact field at x [n, >= dim] -> [dim + 2, n]: the steady-state
    source the reference's ArbitrarySource adds at each cell centroid (arbitrarySource.cpp:15-20)."""
    dim = c.dim
    F, vel, T = _flux_parts(c, x)
    n = x.shape[0]
    S = np.zeros((dim + 2, n))
    for comp in range(dim + 2):
        for d in range(dim):
            S[comp] += F[comp][d].g[:, d]
    if c.mu == 0.0 and c.k == 0.0:
        return S
    # viscous stress tau_cd = mu (du_c/dx_d + du_d/dx_c - 2/3 delta_cd div u) and heat flux k grad T
    div = sum(vel[d].g[:, d] for d in range(dim))
    ddiv = [sum(vel[e].h[:, e, d] for e in range(dim)) for d in range(dim)]  # d(div u)/dx_d
    tau = [[c.mu * (vel[a].g[:, b] + vel[b].g[:, a] - (2.0 / 3.0) * (a == b) * div) for b in range(dim)]
           for a in range(dim)]
    dtau = [[c.mu * (vel[a].h[:, b, b] + vel[b].h[:, a, b] - (2.0 / 3.0) * (a == b) * ddiv[b]) for b in range(dim)]
            for a in range(dim)]  # d tau_ab / dx_b
    for a in range(dim):
        S[2 + a] -= sum(dtau[a][b] for b in range(dim))
    # energy: div (tau u + k grad T)
    e = np.zeros(n)
    for b in range(dim):
        for a in range(dim):
            e += dtau[a][b] * vel[a].v + tau[a][b] * vel[a].g[:, b]
        e += c.k * T.h[:, b, b]
    S[1] -= e
    return S


def source_fd(c: MmsConstants, x: np.ndarray, h: float = 1e-3) -> np.ndarray:
    """The same source by fourth-order central differences of the flux field evaluated from values only (an
    independent check of the jet arithmetic)."""
    dim = c.dim
    n = x.shape[0]
    x3 = np.zeros((n, 3))
    x3[:, : x.shape[1]] = x

    def flux_values(xx):
        F, vel, T = _flux_parts(c, xx)
        out = np.stack([np.stack([F[comp][d].v for d in range(dim)]) for comp in range(dim + 2)])  # [comp][d][n]
        if c.mu or c.k:
            # viscous fluxes from the jets' first derivatives (values of tau and grad T at xx)
            div = sum(vel[d].g[:, d] for d in range(dim))
            for a in range(dim):
                for b in range(dim):
                    tau = c.mu * (vel[a].g[:, b] + vel[b].g[:, a] - (2.0 / 3.0) * (a == b) * div)
                    out[2 + a, b] -= tau
                    out[1, b] -= tau * vel[a].v
            for b in range(dim):
                out[1, b] -= c.k * T.g[:, b]
        return out

    S = np.zeros((dim + 2, n))
    for d in range(dim):
        e = np.zeros(3)
        e[d] = h
        fp1, fm1 = flux_values(x3 + e), flux_values(x3 - e)
        fp2, fm2 = flux_values(x3 + 2 * e), flux_values(x3 - 2 * e)
        S += ((8.0 * (fp1 - fm1) - (fp2 - fm2)) / (12.0 * h))[:, d]
    return S


# ------------------------------------------------------------------------------------------- the rate table
LOW = dict(rho=Phi(1.0, 0.15, -0.1, 0.0, 1.0, 0.5, 0.0), u=Phi(70, 5, -7, 0.0, 1.5, 0.6, 0.0),
           v=Phi(90, -15, -8.5, 0.0, 0.5, 2.0 / 3.0, 0.0), p=Phi(1e5, 0.2e5, 0.5e5, 0.0, 2.0, 1.0, 0.0))
HIGH = dict(rho=Phi(1.0, 0.15, -0.1, 0.0, 1.0, 0.5, 0.0), u=Phi(800, 50, -30.0, 0.0, 1.5, 0.6, 0.0),
            v=Phi(800, -75, 40, 0.0, 0.5, 2.0 / 3.0, 0.0), p=Phi(1e5, 0.2e5, 0.5e5, 0.0, 2.0, 1.0, 0.0))
LOW3D = dict(rho=Phi(1.0, 0.15, -0.1, 0.0, 1.0, 0.5, 0.4), u=Phi(70, 5, -7, 5, 1.5, 0.6, 0.5),
             v=Phi(90, -15, -8.5, 6.5, 0.5, 2.0 / 3.0, 0.6), w=Phi(80, -25, 8.2, -10, 0.75, 0.2, 0.7),
             p=Phi(1e5, 0.2e5, 0.5e5, 0.4e5, 2.0, 1.0, 0.8))


@dataclass
class MmsRow:
    """One row of compressibleFlowMmsSourceTests.cpp:656-895."""

    name: str
    flux: str
    consts: MmsConstants
    nx0: int
    levels: int
    l2: tuple
    linf: tuple
    recon: str = "upwind"
    limiter: str = "none"
    minf: float = 0.0
    source_line: int = 0


def _c(base, dim=2, mu=0.0, k=0.0):
    return MmsConstants(dim=dim, mu=mu, k=k, **base)


ROWS = [
    MmsRow("low speed average", "average", _c(LOW), 4, 4, (2, 2, 2, 2), (1.9, 1.8, 1.8, 1.8), source_line=658),
    MmsRow("high speed average", "average", _c(HIGH), 4, 4, (2, 2, 2, 2), (1.9, 1.8, 1.8, 1.8), source_line=676),
    MmsRow("low speed ausm", "ausm", _c(LOW), 16, 4, (1.0, 1.0, 1.4, 1.0), (1.0, 1.0, 1.4, 1.0), source_line=695),
    MmsRow("high speed ausm", "ausm", _c(HIGH), 16, 4, (1.0, 1.0, 1.0, 1.0), (1.0, 1.0, 1.0, 1.0), source_line=713),
    MmsRow("low speed ausmpup", "ausmpup", _c(LOW), 16, 4, (1.0, 1.0, 1.2, 1.0), (1.0, 1.0, 1.2, 1.0), minf=0.3,
           source_line=731),
    MmsRow("high speed ausmpup", "ausmpup", _c(HIGH), 16, 4, (1.0, 1.0, 1.0, 1.0), (1.0, 1.0, 1.0, 1.0), minf=0.3,
           source_line=749),
    # petscfv_type leastsquares with PETSc's default limiter (PetscLimiterSetFromOptions defaults to sin)
    MmsRow("low speed ausm leastsquares", "ausm", _c(LOW), 16, 4, (1.5, 1.5, 1.5, 1.5), (1.0, 1.0, 1.0, 1.0),
           recon="leastsquares", limiter="sin", source_line=767),
    MmsRow("high speed ausm leastsquares", "ausm", _c(HIGH), 16, 4, (1.5, 1.5, 1.5, 1.5), (1.0, 0.5, 1.0, 1.0),
           recon="leastsquares", limiter="sin", source_line=786),
    MmsRow("low speed average with conduction", "average", _c(LOW, k=1000.0), 4, 4, (2, 2, 2, 2),
           (1.9, 1.8, 1.8, 1.8), source_line=805),
    MmsRow("high speed average with conduction", "average", _c(HIGH, k=1000.0), 4, 4, (2, 2, 2, 2),
           (1.9, 1.8, 1.8, 1.8), source_line=823),
    MmsRow("low speed average with conduction and diffusion", "average", _c(LOW, mu=300.0, k=1000.0), 4, 4,
           (2, 2, 2, 2.2), (1.9, 1.8, 1.8, 2.0), source_line=841),
    MmsRow("high speed average with conduction and diffusion", "average", _c(HIGH, mu=300.0, k=1000.0), 4, 4,
           (2, 2, 2, 2.0), (1.9, 2.0, 1.8, 1.8), source_line=859),
    MmsRow("low speed average with conduction and diffusion 3D", "average", _c(LOW3D, dim=3, mu=300.0, k=1000.0), 10,
           2, (2.1, 2.2, 2.2, 2.2, 2.0), (1.9, 2.2, 2.0, 2.2, 2.0), source_line=877),
]
RATE_TOL = 0.2  # ASSERT_NEAR(slope, expected, 0.2), :643-644


def ghost_function(c: MmsConstants):
    """The test's boundary function (PhysicsBoundary_Euler, :142-155): the exact state at face centroid + normal / 2
    (the area-weighted normal; 'assuming the cells are square')."""

    def fn(time, centroid, normal, xI):
        x = centroid[:, : c.dim] + 0.5 * normal[:, : c.dim]
        return exact_conserved(c, x).T

    return fn


def central_third(mesh, L: float) -> np.ndarray:
    """Interior cells whose centroid lies in [L/3, 2L/3] in every direction (ComputeRHS, :454-529)."""
    x = mesh.arrays.cell_centroid[: mesh.n_interior, : mesh.dim]
    return np.all((x >= L / 3.0) & (x <= 2.0 * L / 3.0), axis=1)


def residual_norms(R: np.ndarray, sel: np.ndarray):
    """Per-component L2 (root mean square over the selected cells) and L-infinity norms."""
    r = R[:, : sel.size][:, sel]
    return np.sqrt(np.mean(r * r, axis=1)), np.abs(r).max(axis=1)


def slopes(hs, norms):
    """Least-squares slope of log10(norm) against log10(h) per component (PetscLinearRegression)."""
    lh = np.log10(np.asarray(hs))
    out = []
    for comp in range(norms.shape[1]):
        ln = np.log10(norms[:, comp])
        out.append(np.polyfit(lh, ln, 1)[0])
    return np.array(out)
