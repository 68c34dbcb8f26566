"""CPU baseline of the benchmark workload on the oracle (TEST / MEASUREMENT INFRASTRUCTURE, never the product).

The reference runs one MPI rank per core, each on its own subdomain, with the per-face loop structure the oracle
keeps (cellInterpolant.cpp:499-639, faceInterpolant.cpp:262-448).  This module times that structure the same way:
P worker processes (one per core), each evaluating the Navier-Stokes RHS (LSQ + Barth-Jespersen + AUSM, Taylor-Green
state) on its own n^3 periodic box for about `seconds`; the throughput is the sum over workers.  It runs as its own
process (``python -m oracle.cpu_baseline``) so that the workers never share a GPU context.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import multiprocessing as mp
import os
import platform
import time


def _worker(args):
    n, seconds, mu, k, mach = args
    import numpy as np

    from ablate_amd import synthetic
    from ablate_amd.mesh import BoxMesh
    from oracle import binding as orc

    mesh = BoxMesh([n, n, n], [0, 0, 0], [1, 1, 1], [True] * 3, ordering="natural")
    U = synthetic.taylor_green(mesh, mach)
    phys = orc.physics(1.4, 287.0, 0, "ausm", "leastsquares", "barthjespersen", 1e10, k, mu, 0.0)
    ctx = orc.Context(mesh.arrays, phys)
    X = np.ascontiguousarray(U.T)
    F = np.zeros_like(X)
    ctx.rhs_aos(X, F)
    calls, t0 = 0, time.perf_counter()
    while True:
        F[:] = 0.0
        ctx.rhs_aos(X, F)
        calls += 1
        el = time.perf_counter() - t0
        if el >= seconds:
            break
    ctx.close()
    return mesh.n_interior, calls, el


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=32)
    ap.add_argument("--seconds", type=float, default=10.0)
    ap.add_argument("--workers", type=int, default=0, help="0 = one per visible core, at most 16")
    ap.add_argument("--mu", type=float, default=1.8e-5)
    ap.add_argument("--k", type=float, default=1.8e-5 * 1.4 * 287.0 / 0.4 / 0.72)
    ap.add_argument("--mach", type=float, default=0.3)
    a = ap.parse_args()
    P = a.workers or min(16, os.cpu_count() or 1)
    job = (a.n, a.seconds, a.mu, a.k, a.mach)
    single = _worker(job)
    with mp.get_context("fork").Pool(P) as pool:
        res = pool.map(_worker, [job] * P)
    rate_1 = single[0] * single[1] / single[2] / 1e6
    rate_p = sum(c * k / e for c, k, e in res) / 1e6
    print(json.dumps({
        "value": round(rate_p, 4),
        "unit": "Mcell/s",
        "cores": P,
        "kind": "port",
        "single_core": round(rate_1, 4),
        "sample": f"C oracle (reference loop structure) on the same NS/LSQ/BJ/AUSM workload, {P} processes (one per "
                  f"core, like one MPI rank per core) each on its own {a.n}^3 periodic box for {a.seconds:.0f} s "
                  f"({sum(r[1] for r in res)} RHS calls in total), aggregate throughput; single process alone: "
                  f"{rate_1:.3f} Mcell/s; host {platform.processor() or 'x86_64'}, {os.cpu_count()} logical cores "
                  f"visible",
    }))


if __name__ == "__main__":
    main()
