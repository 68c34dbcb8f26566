"""CPU baseline of the benchmark workloads on the oracle (TEST / MEASUREMENT INFRASTRUCTURE, never the product).

The reference runs one MPI rank per core, each on its own subdomain, with the per-face loop structure the oracle
keeps (cellInterpolant.cpp:499-639, faceInterpolant.cpp:262-448).  This module times that structure the same way:
P worker processes (one per core), each evaluating the RHS of a workload on its own sample mesh for about
`seconds`; the throughput is the sum over workers.  Workloads (physics of the GPU legs, sample sizes stated):
  ns_hex   Navier-Stokes, LSQ + Barth-Jespersen, AUSM, Taylor-Green state, n^3 periodic hexes (the headline physics)
  euler2d  first-order Euler, AUSM, 64 x 64 periodic quads, random state (the 64 x 64 case at full size)
  species  first-order Euler + 10 species, AUSM, inflow x- / outflow x+, periodic y, z, n^3 hexes
  tets     Navier-Stokes + 10 species (D = 2e-5), LSQ + Barth-Jespersen, Rusanov, outflow ghosts, jittered tets
It runs as its own process (``python -m oracle.cpu_baseline``) so that the workers never share a GPU context.
Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import multiprocessing as mp
import os
import platform
import time

WORKLOADS = ("ns_hex", "euler2d", "species", "tets")


def _setup(workload, n, mu, k, mach):
    """(mesh, state, oracle physics, boundary slots, sample description) of one worker's sample."""
    from ablate_amd import synthetic
    from ablate_amd.mesh import BoxMesh, TetBoxMesh
    from oracle import binding as orc

    if workload == "ns_hex":
        mesh = BoxMesh([n, n, n], [0, 0, 0], [1, 1, 1], [True] * 3, ordering="natural")
        U = synthetic.taylor_green(mesh, mach)
        phys = orc.physics(1.4, 287.0, 0, "ausm", "leastsquares", "barthjespersen", 1e10, k, mu, 0.0)
        return mesh, U, phys, None, f"{n}^3 periodic hexes"
    if workload == "euler2d":
        mesh = BoxMesh([64, 64], [0, 0], [1, 1], [True, True], ordering="natural")
        U = synthetic.random_euler_2d(mesh)
        return mesh, U, orc.physics(1.4, 287.0, 0, "ausm", "upwind"), None, "64 x 64 periodic quads (full size)"
    ns = 10
    if workload == "species":
        mesh = BoxMesh([n, n, n], [0, 0, 0], [1, 1, 1], [False, True, True], ordering="natural")
        U = synthetic.taylor_green(mesh, 0.3, n_species=ns, phase=(0.37, 0.21, 0.53))
        phys = orc.physics(1.4, 287.0, ns, "ausm", "upwind")
        bcs = {0: {"kind": orc.BC_INFLOW, "rho": 1.2, "vel": [100.0, 0.0, 0.0], "T": 300.0, "Yi": [1.0 / ns] * ns},
               1: {"kind": orc.BC_OUTFLOW}}
        return mesh, U, phys, bcs, f"{n}^3 hexes, inflow x- / outflow x+"
    if workload == "tets":
        nt = max(2, int(round(n / 6 ** (1.0 / 3.0))))
        mesh = TetBoxMesh(nt, 0.2)
        U = synthetic.taylor_green(mesh, 0.3, n_species=ns)
        phys = orc.physics(1.4, 287.0, ns, "rusanov", "leastsquares", "barthjespersen", 1e10, k, mu, 2e-5)
        bcs = {s: {"kind": orc.BC_OUTFLOW} for s in range(6)}
        return mesh, U, phys, bcs, f"{nt}^3 x 6 jittered tets ({mesh.n_interior} cells), outflow ghosts"
    raise ValueError(workload)


def _worker(args):
    workload, n, seconds, mu, k, mach = args
    import numpy as np

    from oracle import binding as orc

    mesh, U, phys, bcs, _ = _setup(workload, n, mu, k, mach)
    ctx = orc.Context(mesh.arrays, phys)
    X = np.ascontiguousarray(U.T)
    F = np.zeros_like(X)
    ctx.rhs_aos(X, F, bcs)
    calls, t0 = 0, time.perf_counter()
    while True:
        F[:] = 0.0
        ctx.rhs_aos(X, F, bcs)
        calls += 1
        el = time.perf_counter() - t0
        if el >= seconds:
            break
    ctx.close()
    return mesh.n_interior, calls, el


def measure(workload, n, seconds, P, mu, k, mach):
    job = (workload, n, seconds, mu, k, mach)
    single = _worker(job)
    with mp.get_context("fork").Pool(P) as pool:
        res = pool.map(_worker, [job] * P)
    rate_1 = single[0] * single[1] / single[2] / 1e6
    rate_p = sum(c * q / e for c, q, e in res) / 1e6
    desc = _setup(workload, n, mu, k, mach)[4]
    return {
        "value": round(rate_p, 4),
        "unit": "Mcell/s",
        "cores": P,
        "kind": "port",
        "single_core": round(rate_1, 4),
        "sample": f"C oracle (reference loop structure), {workload} physics, {P} processes (one per core, like one "
                  f"MPI rank per core) each on its own {desc} for {seconds:.0f} s ({sum(r[1] for r in res)} RHS calls "
                  f"in total), aggregate throughput; single process alone: {rate_1:.3f} Mcell/s; host "
                  f"{platform.processor() or 'x86_64'}, {os.cpu_count()} logical cores visible",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="ns_hex", choices=WORKLOADS + ("all",))
    ap.add_argument("--n", type=int, default=32, help="cells per direction of the hex samples (tets: about n^3 cells)")
    ap.add_argument("--seconds", type=float, default=10.0)
    ap.add_argument("--other-seconds", type=float, default=4.0, help="time per phase of the other workloads (all)")
    ap.add_argument("--workers", type=int, default=0, help="0 = one per visible core, at most 16")
    ap.add_argument("--mu", type=float, default=1.8e-5)
    ap.add_argument("--k", type=float, default=1.8e-5 * 1.4 * 287.0 / 0.4 / 0.72)
    ap.add_argument("--mach", type=float, default=0.3)
    a = ap.parse_args()
    P = a.workers or min(16, os.cpu_count() or 1)
    if a.workload != "all":
        print(json.dumps(measure(a.workload, a.n, a.seconds, P, a.mu, a.k, a.mach)))
        return
    out = measure("ns_hex", a.n, a.seconds, P, a.mu, a.k, a.mach)
    out["other_workloads"] = {w: measure(w, a.n, a.other_seconds, P, a.mu, a.k, a.mach)
                              for w in ("euler2d", "species", "tets")}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
