"""Seeded synthetic states for the FV RHS workloads (all randomness from one counter-based generator).

Every random number is splitmix64(seed, stream, counter) with the cell's canonical index as the counter (the
mesh producer's own cell number, ``mesh.source_cells()``: the lexicographic box index, the tet generation index,
the DMPlex point or the Gmsh element), so every cell ordering of a mesh, CPU and GPU runs and runs of different
processes see the same state on the same physical cell.  States are structure-of-arrays ``U[comp, cell]`` in the
reference's component order (rho, rhoE, rhoU, rhoV[, rhoW], rhoY_0 ..), including the ghost rows (zero until the
RHS fills them).
"""
from __future__ import annotations

import numpy as np

DEFAULT_SEED = 20261202
_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _splitmix64(x: np.ndarray) -> np.ndarray:
    x = (x + np.uint64(0x9E3779B97F4A7C15)) & _M64
    x = ((x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & _M64
    x = ((x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & _M64
    return x ^ (x >> np.uint64(31))


def _counters(cells) -> np.ndarray:
    """Counter array: an int n means counters 0..n-1, an array gives the counters themselves (canonical cell ids)."""
    if np.isscalar(cells):
        return np.arange(int(cells), dtype=np.uint64)
    return np.asarray(cells, dtype=np.int64).astype(np.uint64)


def uniform01(seed: int, stream: int, cells) -> np.ndarray:
    """U[0,1) of a stream at the given counters (53-bit mantissa)."""
    with np.errstate(over="ignore"):
        key = _splitmix64(np.uint64(seed) ^ _splitmix64(np.asarray([stream], dtype=np.uint64)))[0]
        r = _splitmix64(key ^ _splitmix64(_counters(cells)))
    return (r >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def uniform_pm1(seed: int, stream: int, cells) -> np.ndarray:
    return 2.0 * uniform01(seed, stream, cells) - 1.0


def normal(seed: int, stream: int, cells) -> np.ndarray:
    """N(0,1) by Box-Muller on two counter streams."""
    u1 = uniform01(seed, 2 * stream + 1000, cells)
    u2 = uniform01(seed, 2 * stream + 1001, cells)
    u1 = np.maximum(u1, 1e-300)
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)


def conserved_from_primitive(rho, vel, p, gamma):
    """rho, velocity [dim, n], pressure -> euler components [dim + 2, n]."""
    dim = vel.shape[0]
    U = np.empty((dim + 2, rho.shape[0]))
    U[0] = rho
    U[1] = p / (gamma - 1.0) + 0.5 * rho * np.sum(vel * vel, axis=0)
    U[2:] = rho * vel
    return U


def random_euler_2d(mesh, gamma=1.4, seed=DEFAULT_SEED):
    """64x64-style seeded Euler state: rho = 1 +/- 10 %, u, v ~ U(-50, 50) m/s, p = 1e5 +/- 10 % Pa."""
    N, NI = mesh.n_cells, mesh.n_interior
    cells = mesh.source_cells()
    rho = 1.0 * (1.0 + 0.1 * uniform_pm1(seed, 0, cells))
    vel = np.stack([50.0 * uniform_pm1(seed, 1, cells), 50.0 * uniform_pm1(seed, 2, cells)])
    p = 1e5 * (1.0 + 0.1 * uniform_pm1(seed, 3, cells))
    U = np.zeros((4, N))
    U[:, :NI] = conserved_from_primitive(rho, vel, p, gamma)
    return U


def taylor_green(mesh, mach=0.3, gamma=1.4, R=287.0, p0=1e5, T0=300.0, noise=0.01, seed=DEFAULT_SEED, n_species=0,
                 phase=(0.0, 0.0, 0.0)):
    """Taylor-Green vortex on [0, 2pi)^3 mapped to the mesh box, at Mach ``mach``; rho, rhoU, rhoE then scaled by
    (1 + noise * U(-1, 1)) independently; optional Dirichlet(1) species mass fractions.  ``phase`` shifts the field
    (radians per axis): a non-zero phase keeps cell centres off the vortex symmetry planes, where a velocity
    component vanishes exactly and AUSM's upwind switch sits at the evaluation point (used for derivative checks)."""
    N, NI, dim = mesh.n_cells, mesh.n_interior, mesh.dim
    x = mesh.arrays.cell_centroid[:NI]
    lo = np.asarray(getattr(mesh, "lower", [0.0, 0.0, 0.0]) + [0.0] * (3 - dim))[:3]
    hi = np.asarray(getattr(mesh, "upper", [1.0, 1.0, 1.0]) + [1.0] * (3 - dim))[:3]
    xt = 2.0 * np.pi * (x - lo) / (hi - lo) + np.asarray(phase, dtype=np.float64)[None, :]
    rho0 = p0 / (R * T0)
    V0 = mach * np.sqrt(gamma * R * T0)
    sx, sy, sz = np.sin(xt[:, 0]), np.sin(xt[:, 1]), np.sin(xt[:, 2])
    cx, cy, cz = np.cos(xt[:, 0]), np.cos(xt[:, 1]), np.cos(xt[:, 2])
    vel = np.zeros((dim, NI))
    vel[0] = V0 * sx * cy * cz
    vel[1] = -V0 * cx * sy * cz
    p = p0 + rho0 * V0 * V0 / 16.0 * (np.cos(2 * xt[:, 0]) + np.cos(2 * xt[:, 1])) * (np.cos(2 * xt[:, 2]) + 2.0)
    rho = np.full(NI, rho0)
    Ue = conserved_from_primitive(rho, vel, p, gamma)
    cells = mesh.source_cells()
    Ue[0] *= 1.0 + noise * uniform_pm1(seed, 10, cells)
    for d in range(dim):
        Ue[2 + d] *= 1.0 + noise * uniform_pm1(seed, 11 + d, cells)
    Ue[1] *= 1.0 + noise * uniform_pm1(seed, 14, cells)
    U = np.zeros((dim + 2 + n_species, N))
    U[: dim + 2, :NI] = Ue
    if n_species:
        U[dim + 2 :, :NI] = Ue[0] * dirichlet_mass_fractions(n_species, cells, seed)
    return U


def dirichlet_mass_fractions(ns: int, cells, seed=DEFAULT_SEED) -> np.ndarray:
    """Y ~ Dirichlet(1): Y_i = -ln U_i / sum_j -ln U_j  ->  [ns, n] (cells: a count or canonical cell ids)."""
    e = np.stack([-np.log(np.maximum(uniform01(seed, 100 + i, cells), 1e-300)) for i in range(ns)])
    return e / e.sum(axis=0, keepdims=True)


def adjoint_seed(nc: int, n_cells: int, n_interior: int, seed=DEFAULT_SEED, cells=None) -> np.ndarray:
    """Seed vector Rbar ~ N(0, 1) on interior rows (counters: the canonical cell ids when given), zero on ghost
    rows."""
    out = np.zeros((nc, n_cells))
    ctr = n_interior if cells is None else cells
    for p in range(nc):
        out[p, :n_interior] = normal(seed, 200 + p, ctr)
    return out
