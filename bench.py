"""Benchmark of the MI355X finite-volume RHS (one step = one abl_fv_rhs evaluation on a resident state).

Default workload: Navier-Stokes RHS in FP64 on a 256^3 periodic hex mesh (16.8M cells, 50.3M faces), gamma = 1.4,
mu = 1.8e-5, k = mu c_p / Pr (Pr = 0.72), least-squares gradients with the Barth-Jespersen limiter, AUSM fluxes,
seeded Taylor-Green state at Mach 0.3 with 1 % perturbations on rho, rhoU, rhoE.  Prints ONE JSON line.

Roofline: the RHS is a sparse gather/stencil with no dense contraction, so its bound is HBM bandwidth.  The
achieved figure divides the algorithmic bytes of one RHS evaluation (state read + residual write + the mesh data
one evaluation must read, abl_fv_get_info().alg_bytes_per_cell x cells) by the event-timed device duration of the
RHS kernels in the timed region.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0
CONFIGS = {
    "c2": dict(workload="ns_fp64_256cube_lsq_bj_ausm", n=256, form="fp64", mach=0.3),
    "c3": dict(workload="ns_mixed_256cube_lsq_bj_ausm", n=256, form="mixed", mach=0.3),
    "c3_lowmach": dict(workload="ns_mixed_256cube_lsq_bj_ausm_mach0.01", n=256, form="mixed", mach=0.01),
}


def parse():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="c2", choices=sorted(CONFIGS))
    ap.add_argument("--n", type=int, default=None, help="cells per direction (default: the workload's 256)")
    ap.add_argument("--strategy", default="auto", choices=["auto", "recompute", "face_flux", "tile", "fused", "march"],
                    help="evaluation strategy of the headline leg (auto: the library's default choice)")
    ap.add_argument("--ordering", default="auto", choices=["auto", "brick", "column", "natural"],
                    help="cell order of the headline mesh (auto: column for the march strategy, else brick)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision leg")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint (VJP) leg")
    ap.add_argument("--adjoint-n", type=int, default=128, help="cells per direction of the adjoint workload")
    ap.add_argument("--no-tets", action="store_true", help="skip the tetrahedral NS + species leg")
    ap.add_argument("--no-c1", action="store_true", help="skip the 64 x 64 first-order Euler leg (eager and HIP graph)")
    ap.add_argument("--tets-n", type=int, default=111, help="cubes per direction of the tetrahedral workload")
    ap.add_argument("--cpu-n", type=int, default=32, help="cells per direction of each CPU-baseline process's sample")
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="CPU-baseline time per phase")
    ap.add_argument("--copy-gib", type=float, default=2.0,
                    help="size of the plain device copy timed before the RHS (HBM reference and counter calibration; "
                         "0 = skip)")
    return ap.parse_args()


def copy_probe(gib, reps=5):
    """Times a plain device-to-device copy of `gib` GiB (torch copy_, one read + one write per byte) on the current
    stream.  Under a hardware-counter profiler it is the calibration kernel for FETCH_SIZE / WRITE_SIZE (its
    algorithmic traffic is exactly `gib` read + `gib` written); its rate is below what the memory system delivers
    (see stream_probe)."""
    import torch

    n = int(gib * (1 << 30)) // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        b.copy_(a)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    nbytes = 2 * n * 8
    del a, b
    torch.cuda.empty_cache()
    return {"bytes_read": n * 8, "bytes_written": n * 8, "ms": round(ms, 4),
            "GB_per_s": round(nbytes / (ms * 1e-3) / 1e9, 1)}


def stream_probe(gib, reps=10):
    """Achievable HBM bandwidth: the library's streaming copy and streaming read (abl_hbm_probe, 16-byte loads per
    lane) over `gib` GiB, after warm-up passes."""
    import ctypes

    from ablate_amd import _lib

    out = {"bytes": int(gib * (1 << 30))}
    for kind, name in ((0, "copy"), (1, "read")):
        gbs, ms = ctypes.c_double(), ctypes.c_double()
        _lib.check(_lib.lib().abl_hbm_probe(kind, out["bytes"], reps, ctypes.byref(gbs), ctypes.byref(ms)), None,
                   "abl_hbm_probe")
        out[f"{name}_GB_per_s"] = round(gbs.value, 1)
    return out


KERNELS = {"face_flux": ["k_gradients_w4", "k_face_flux", "k_gather"], "recompute": ["k_gradients_w4", "k_flux"],
           "tile": ["k_gradients_w4", "k_tile_flux"], "fused": ["k_fused"], "march": ["k_march"]}


def recorded_traffic(workload, n, strategy):
    """HBM bytes per call from hardware counters: the committed record of profiles/traffic.json for this workload,
    measured on the same device sources (FETCH_SIZE and WRITE_SIZE per dispatch summed over the call's kernels, each
    scaled by what the counters read for the streaming copy of abl_hbm_probe; see ablate_amd/rocprof.py).  None
    when no record matches the sources that built the library."""
    from ablate_amd import rocprof

    rec = rocprof.lookup(workload, n, strategy)
    return None if rec is None else rec["traffic_bytes"]


def measured_traffic(cfg, n, strategy):
    """HBM bytes of one RHS of a headline-mesh workload (gradient and flux kernels), or None."""
    return recorded_traffic(cfg["workload"], n, strategy)


def tets_traffic(n, strategy, form="fp64"):
    """HBM bytes of one RHS of the tets leg (ghost fill, gradient, species-gradient, face-flux and gather kernels;
    form "fp64" or "mixed"), or of one FP64 VJP call (form "adjoint": ghost fill, the forward gradient passes and
    the adjoint passes), or None."""
    if form == "adjoint":
        return recorded_traffic(f"vjp_fp64_tets{n}_ns_10species_lsq_bj_rusanov", n, "adjoint")
    return recorded_traffic(f"ns_10species_{form}_tets{n}_lsq_bj_rusanov", n, strategy)


def ns_processes(fv, mu, k):
    eos = fv.PerfectGas({"gamma": 1.4, "Rgas": 287.0})
    return [fv.NavierStokesTransport({"cfl": 0.5}, eos, fv.Ausm(), fv.Constant(k=k, mu=mu))]


def cpu_baseline(n, seconds, mu, k, mach):
    """Times the CPU oracle (a C restatement of the reference loops) on the host cores: one process per core, each on
    its own n^3 sample of the same workload (the reference runs one MPI rank per core), plus one process alone; then
    the physics of the other legs (64 x 64 first-order Euler, first-order Euler + 10 species with inflow, tets with
    NS + 10 species) under `other_workloads`.  Runs as a child process (oracle/cpu_baseline.py) so that its workers
    never share this process's GPU context."""
    import subprocess

    root = os.path.dirname(os.path.abspath(__file__))
    cmd = [sys.executable, "-m", "oracle.cpu_baseline", "--workload", "all", "--n", str(n), "--seconds", str(seconds),
           "--other-seconds", str(max(1.0, 0.4 * seconds)), "--mu", repr(mu), "--k", repr(k), "--mach", repr(mach)]
    out = subprocess.run(cmd, cwd=root, capture_output=True, text=True, timeout=max(240.0, 12 * seconds + 120))
    if out.returncode != 0:
        return {"error": f"cpu baseline failed ({out.returncode}): {out.stderr[-400:]}"}
    return json.loads(out.stdout.strip().splitlines()[-1])


def progress(msg):
    """One line on stderr per finished leg (the JSON result stays the only line on stdout)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def per_field_rel(a, b):
    return [float(x) for x in ((a - b).norm(dim=1) / b.norm(dim=1)).tolist()]


def mixed_leg(fv, synthetic, mesh, mu, k, R64, U, mach, args, fp64_ms):
    """Mixed form on the same mesh: timed like the FP64 leg; per-field error against the FP64 GPU residual at
    Mach 0.3 and at the low-Mach variant (Mach 0.01)."""
    import torch

    solver = fv.FiniteVolumeSolver(mesh, ns_processes(fv, mu, k),
                                   {"petscfv_type": "leastsquares", "petsclimiter_type": "barthjespersen"},
                                   form="mixed")
    Rm = torch.zeros_like(U)
    for _ in range(args.warmup):
        solver.compute_rhs(0.0, U, Rm)
    torch.cuda.synchronize()
    solver.set_profiling(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        solver.compute_rhs(0.0, U, Rm)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    solver.set_profiling(False)
    phases = solver.phase_times(reset=True)
    err = per_field_rel(Rm, R64)
    # low-Mach variant: FP64 reference and mixed on the Mach 0.01 state
    U2 = torch.tensor(synthetic.taylor_green(mesh, 0.01), dtype=torch.float64, device="cuda")
    ref = fv.FiniteVolumeSolver(mesh, ns_processes(fv, mu, k),
                                {"petscfv_type": "leastsquares", "petsclimiter_type": "barthjespersen"})
    R2d, R2m = torch.zeros_like(U2), torch.zeros_like(U2)
    ref.compute_rhs(0.0, U2, R2d)
    solver.compute_rhs(0.0, U2, R2m)
    torch.cuda.synchronize()
    err_low = per_field_rel(R2m, R2d)
    del ref
    cells = mesh.n_interior
    strategy = solver.strategy()
    return {
        "value": round(cells / (ms * 1e-3) / 1e6, 3),
        "unit": "Mcell/s",
        "ms_per_step": round(ms, 4),
        "speedup_vs_fp64": round(fp64_ms / ms, 3),
        "strategy": strategy,
        "traffic": measured_traffic(CONFIGS["c3"], args.n or CONFIGS["c3"]["n"], strategy),
        "phases_ms": {name: round(v / max(c, 1), 4) for name, (v, c) in phases.items() if c},
        "rel_err_vs_fp64_mach0.3": [float(f"{e:.3e}") for e in err],
        "rel_err_vs_fp64_mach0.01": [float(f"{e:.3e}") for e in err_low],
        "error_norm": "per conserved field ||R_mixed - R_fp64|| / ||R_fp64|| over all cells (rho, rhoE, rhoU, rhoV, rhoW)",
    }


def tets_leg(fv, synthetic, args):
    """C5: jittered tetrahedral unit cube (n^3 cubes x 6 Kuhn tets, interior nodes moved by up to 20 % of h, RCM
    order), Navier-Stokes + 10 species (D = 2e-5), LSQ + Barth-Jespersen, Rusanov, zero-gradient ghost cells on
    all six sides; Taylor-Green state at Mach 0.3 with Dirichlet(1) mass fractions.  FP64 forward timed like the
    headline leg, the mixed forward with its error against FP64, and the FP64 adjoint (seed N(0, 1)) with a
    directional check against central differences (h = 1e-6) of the GPU forward."""
    import torch

    from ablate_amd.mesh import TetBoxMesh

    n, ns = args.tets_n, 10
    t0 = time.perf_counter()
    mesh = TetBoxMesh(n, 0.2)
    species = [f"Y{i}" for i in range(ns)]
    eos = fv.PerfectGas({"gamma": 1.4, "Rgas": 287.0}, species)
    mu = 1.8e-5
    tr = fv.Constant(k=mu * 1.4 * 287.0 / 0.4 / 0.72, mu=mu, diff=2e-5)
    procs = [fv.NavierStokesTransport({"cfl": 0.5}, eos, fv.Rusanov(), tr), fv.SpeciesTransport(eos, fv.Rusanov(), tr)]
    bcs = [fv.OutflowGhost(list(range(6)))]
    solver = fv.FiniteVolumeSolver(mesh, procs, {"petscfv_type": "leastsquares",
                                                 "petsclimiter_type": "barthjespersen"}, bcs)
    U = torch.tensor(synthetic.taylor_green(mesh, 0.3, n_species=ns), dtype=torch.float64, device="cuda")
    R = torch.zeros_like(U)
    setup_s = time.perf_counter() - t0
    info = solver.info()
    strategy = solver.strategy()
    for _ in range(args.warmup):
        solver.compute_rhs(0.0, U, R)
    torch.cuda.synchronize()
    solver.phase_times(reset=True)
    solver.set_profiling(True)
    t1 = time.perf_counter()
    for _ in range(args.steps):
        solver.compute_rhs(0.0, U, R)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t1) * 1e3 / args.steps
    solver.set_profiling(False)
    phases = solver.phase_times(reset=True)
    kernel_ms = sum(v for v, _ in phases.values()) / args.steps
    cells = info.n_interior
    achieved = info.alg_bytes_per_cell * cells / (kernel_ms * 1e-3) / 1e9
    # mixed form on the same workload, error per field against the FP64 residual
    sm = fv.FiniteVolumeSolver(mesh, procs, {"petscfv_type": "leastsquares", "petsclimiter_type": "barthjespersen"},
                               bcs, form="mixed")
    strategy_mixed = sm.strategy()
    Rm = torch.zeros_like(U)
    for _ in range(args.warmup):
        sm.compute_rhs(0.0, U, Rm)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    for _ in range(args.steps):
        sm.compute_rhs(0.0, U, Rm)
    torch.cuda.synchronize()
    ms_mixed = (time.perf_counter() - t2) * 1e3 / args.steps
    err = per_field_rel(Rm, R)
    del sm
    # FP64 adjoint on the same workload (second order: LSQ + Barth-Jespersen, viscous, species diffusion)
    Rb = torch.tensor(synthetic.adjoint_seed(U.shape[0], mesh.n_cells, mesh.n_interior, cells=mesh.source_cells()), device="cuda")
    Ub = torch.empty_like(U)
    for _ in range(args.warmup):
        solver.rhs_vjp(0.0, U, Rb, Ub)
    torch.cuda.synchronize()
    solver.phase_times(reset=True)
    solver.set_profiling(True)
    t3 = time.perf_counter()
    for _ in range(args.steps):
        solver.rhs_vjp(0.0, U, Rb, Ub)
    torch.cuda.synchronize()
    ms_adj = (time.perf_counter() - t3) * 1e3 / args.steps
    solver.set_profiling(False)
    adj_phases = solver.phase_times(reset=True)
    # Full-size checks along directions on which the RHS is affine, so that a central difference is exact up to
    # rounding whatever the limiter and flux branches: scaling the whole state (V = U) leaves velocity, temperature,
    # mass fractions, Mach numbers and every Barth-Jespersen ratio unchanged (the inviscid fluxes and the species
    # diffusion scale with the state, the Navier-Stokes viscous flux does not change), and scaling the species
    # densities alone (V = [0, rhoY]) leaves the euler fluxes unchanged and scales the species fluxes.
    NI = mesh.n_interior
    ne = mesh.dim + 2
    tol = 1e-8
    checks = {}
    for name, V in (("scale_all", U.clone()), ("scale_species", torch.cat([torch.zeros_like(U[:ne]), U[ne:]]))):
        V[:, NI:] = 0.0
        h = 1e-3
        Rp, Rq = torch.zeros_like(U), torch.zeros_like(U)
        solver.compute_rhs(0.0, U + h * V, Rp)
        solver.compute_rhs(0.0, U - h * V, Rq)
        fd = float((Rb * (Rp - Rq)).sum().item() / (2 * h))
        ad = float((Ub * V).sum().item())
        rel = abs(fd - ad) / abs(fd)
        checks[name] = {"fd_central_h1e-3": fd, "vjp": ad, "rel_diff": rel, "tolerance": tol, "pass": rel <= tol}
    # a random direction for information: Barth-Jespersen makes R piecewise smooth, and cells whose limiter branch
    # switches within +-h bias its central difference (the bias shrinks linearly with h)
    g = torch.Generator(device="cuda").manual_seed(9)
    V = torch.zeros_like(U)
    V[:, :NI] = torch.randn(U.shape[0], NI, dtype=torch.float64, device="cuda", generator=g) * U[:, :NI].abs()
    ad = float((Ub * V).sum().item())
    rnd = {"direction": "|U| N(0,1)", "vjp": ad}
    for h in (1e-6, 1e-7):
        Rp, Rq = torch.zeros_like(U), torch.zeros_like(U)
        solver.compute_rhs(0.0, U + h * V, Rp)
        solver.compute_rhs(0.0, U - h * V, Rq)
        fd = float((Rb * (Rp - Rq)).sum().item() / (2 * h))
        rnd[f"rel_diff_h{h:g}"] = abs(fd - ad) / abs(fd)
    checks["random_direction_information"] = rnd
    return {
        "workload": f"ns_10species_fp64_tets{n}_lsq_bj_rusanov",
        "cells": cells,
        "faces": info.n_faces,
        "strategy": strategy,
        "value": round(cells / (ms * 1e-3) / 1e6, 3),
        "unit": "Mcell/s",
        "ms_per_step": round(ms, 4),
        "phases_ms": {name: round(v / max(c, 1), 4) for name, (v, c) in phases.items() if c},
        "alg_bytes_per_cell": round(info.alg_bytes_per_cell, 1),
        "achieved_GB_per_s": round(achieved, 1),
        "frac_of_8TB_s": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": tets_traffic(n, strategy),
        "traffic_GB_per_s": (None if tets_traffic(n, strategy) is None
                             else round(tets_traffic(n, strategy) / (kernel_ms * 1e-3) / 1e9, 1)),
        "setup_s": round(setup_s, 1),
        "finite": bool(torch.isfinite(R).all().item()),
        "mixed": {"value": round(cells / (ms_mixed * 1e-3) / 1e6, 3), "unit": "Mcell/s",
                  "ms_per_step": round(ms_mixed, 4), "speedup_vs_fp64": round(ms / ms_mixed, 3),
                  "strategy": strategy_mixed,
                  "traffic": tets_traffic(n, strategy_mixed, "mixed"),
                  "rel_err_vs_fp64_max": float(f"{max(err):.3e}"),
                  "error_norm": "max over the 15 conserved fields of ||R_mixed - R_fp64|| / ||R_fp64||"},
        "adjoint": {"value": round(cells / (ms_adj * 1e-3) / 1e6, 3), "unit": "Mcell/s", "ms_per_step": round(ms_adj, 4),
                    "adjoint_over_forward": round(ms_adj / ms, 3),
                    "traffic": tets_traffic(n, strategy, "adjoint"),
                    "phases_ms": {name: round(v / max(c, 1), 4) for name, (v, c) in adj_phases.items() if c},
                    "directional_checks": checks,
                    "checks_pass": all(c["pass"] for c in checks.values() if "pass" in c)},
    }


def c1_leg(fv, synthetic, args):
    """The reference's CPU-runnable case on the GPU: first-order Euler RHS (AUSM) on a 64 x 64 periodic quad box,
    rho = 1 +- 10 %, u, v ~ U(-50, 50) m/s, p = 1e5 +- 10 %.  At 4096 cells one RHS is launch-bound: timed eagerly
    (one call after another on the stream) and as replays of a HIP graph holding 50 RHS calls."""
    import torch

    from ablate_amd.mesh import BoxMesh

    mesh = BoxMesh([64, 64], [0, 0], [1, 1], [True, True])
    eos = fv.PerfectGas({"gamma": 1.4, "Rgas": 287.0})
    solver = fv.FiniteVolumeSolver(mesh, [fv.NavierStokesTransport({"cfl": 0.5}, eos, fv.Ausm())],
                                   {"petscfv_type": "upwind"})
    U = torch.tensor(synthetic.random_euler_2d(mesh), dtype=torch.float64, device="cuda")
    R = torch.zeros_like(U)
    stream = torch.cuda.current_stream()
    calls = 200
    for _ in range(10):
        solver.compute_rhs(0.0, U, R)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(calls):
        solver.compute_rhs(0.0, U, R)
    e1.record(stream)
    torch.cuda.synchronize()
    eager_us = e0.elapsed_time(e1) * 1e3 / calls
    per_graph = 50
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        for _ in range(per_graph):
            solver.compute_rhs(0.0, U, R)
    g.replay()
    torch.cuda.synchronize()
    reps = calls // per_graph
    e0.record(stream)
    for _ in range(reps):
        g.replay()
    e1.record(stream)
    torch.cuda.synchronize()
    graph_us = e0.elapsed_time(e1) * 1e3 / (reps * per_graph)
    cells = mesh.n_interior
    return {"workload": "euler_fp64_64x64_first_order_ausm", "cells": cells, "strategy": solver.strategy(),
            "eager_us_per_rhs": round(eager_us, 2), "eager_Mcell_per_s": round(cells / eager_us, 1),
            "graph_us_per_rhs": round(graph_us, 2), "graph_Mcell_per_s": round(cells / graph_us, 1),
            "note": f"graph: {per_graph} RHS calls captured once and replayed; launch-bound at this size"}


def adjoint_leg(fv, synthetic, args):
    """C4: VJP of the first-order FP64 RHS on a 128^3 hex mesh, Euler + 10 species, inflow (rho 1.2, u 100 m/s,
    T 300 K) on x-, outflow on x+, periodic y/z; base flow the Taylor-Green field at Mach 0.3 (phase-shifted off
    its symmetry planes) with Dirichlet(1) mass fractions; seed Rbar ~ N(0, 1).  Timed like the forward legs, plus
    full-size directional-derivative checks: Rbar.(R(U+hV) - R(U-hV))/2h of the GPU FP64 forward along the
    species-density scaling direction (branch-free, pass / fail) and a random direction (information)."""
    import torch

    from ablate_amd.mesh import BoxMesh

    n, ns = args.adjoint_n, 10
    mesh = BoxMesh([n, n, n], [0, 0, 0], [1, 1, 1], [False, True, True], ordering="brick")
    species = [f"Y{i}" for i in range(ns)]
    eos = fv.PerfectGas({"gamma": 1.4, "Rgas": 287.0}, species)
    procs = [fv.NavierStokesTransport({"cfl": 0.5}, eos, fv.Ausm()), fv.SpeciesTransport(eos, fv.Ausm())]
    bcs = [fv.InflowGhost([0], 1.2, [100.0, 0.0, 0.0], 300.0, [1.0 / ns] * ns), fv.OutflowGhost([1])]
    solver = fv.FiniteVolumeSolver(mesh, procs, {"petscfv_type": "upwind"}, bcs)
    U0 = synthetic.taylor_green(mesh, 0.3, n_species=ns, phase=(0.37, 0.21, 0.53))
    U = torch.tensor(U0, dtype=torch.float64, device="cuda")
    Rb = torch.tensor(synthetic.adjoint_seed(U.shape[0], mesh.n_cells, mesh.n_interior, cells=mesh.source_cells()), device="cuda")
    Ub = torch.empty_like(U)
    for _ in range(args.warmup):
        solver.rhs_vjp(0.0, U, Rb, Ub)
    torch.cuda.synchronize()
    solver.set_profiling(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        solver.rhs_vjp(0.0, U, Rb, Ub)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    solver.set_profiling(False)
    phases = solver.phase_times(reset=True)
    theta_bar = solver.rhs_vjp(0.0, U, Rb, Ub)[0]
    # forward timing on the same workload (for the adjoint / forward cost ratio)
    R = torch.zeros_like(U)
    solver.compute_rhs(0.0, U, R)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        solver.compute_rhs(0.0, U, R)
    torch.cuda.synchronize()
    fwd_ms = (time.perf_counter() - t0) * 1e3 / args.steps
    # Directional derivative checks at full size.  Scaling the species densities alone (V = [0, rhoY]) leaves the
    # euler fluxes and every AUSM branch unchanged and scales the species fluxes, so its central difference is exact
    # up to rounding (pass / fail at a stated tolerance); a random direction is reported for information, since the
    # AUSM direction switch and the split polynomials' |M| = 1 seams make the RHS piecewise smooth.
    NI = mesh.n_interior
    ne = mesh.dim + 2
    tol = 1e-8
    checks = {}
    Vs = torch.cat([torch.zeros_like(U[:ne]), U[ne:]])
    Vs[:, NI:] = 0.0
    hs = 1e-3
    Rp, Rm = torch.zeros_like(U), torch.zeros_like(U)
    solver.compute_rhs(0.0, U + hs * Vs, Rp)
    solver.compute_rhs(0.0, U - hs * Vs, Rm)
    fd_s = float((Rb * (Rp - Rm)).sum().item() / (2 * hs))
    ad_s = float((Ub * Vs).sum().item())
    rel_s = abs(fd_s - ad_s) / abs(fd_s)
    checks["scale_species"] = {"fd_central_h1e-3": fd_s, "vjp": ad_s, "rel_diff": rel_s, "tolerance": tol,
                               "pass": rel_s <= tol}
    g = torch.Generator(device="cuda").manual_seed(5)
    V = torch.zeros_like(U)
    scale = U[:, :NI].abs().mean(dim=1, keepdim=True)
    V[:, :NI] = torch.randn(U.shape[0], NI, dtype=torch.float64, device="cuda", generator=g) * scale
    h = 1e-6
    solver.compute_rhs(0.0, U + h * V, Rp)
    solver.compute_rhs(0.0, U - h * V, Rm)
    fd = float((Rb * (Rp - Rm)).sum().item() / (2 * h))
    ad = float((Ub * V).sum().item())
    checks["random_direction_information"] = {"direction": "mean|U| N(0,1)", "fd_central_h1e-6": fd, "vjp": ad,
                                              "rel_diff": abs(fd - ad) / abs(fd)}
    cells = NI
    return {
        "value": round(cells / (ms * 1e-3) / 1e6, 3),
        "unit": "Mcell/s",
        "ms_per_step": round(ms, 4),
        "forward_ms_per_step": round(fwd_ms, 4),
        "adjoint_over_forward": round(ms / fwd_ms, 3),
        "phases_ms": {name: round(v / max(c, 1), 4) for name, (v, c) in phases.items() if c},
        "workload": f"vjp_fp64_{n}cube_euler_10species_inflow_ausm_first_order",
        "cells": cells,
        "traffic": recorded_traffic(f"vjp_fp64_{n}cube_euler_10species_inflow_ausm_first_order", n, "adjoint"),
        "directional_checks": checks,
        "checks_pass": all(c["pass"] for c in checks.values() if "pass" in c),
        "theta_bar_inflow": [float(x) for x in theta_bar],
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "the RHS runs on one GPU; multi-GPU sharding (halo exchange between domain "
                                   "partitions) is designed but not implemented, so --gpus must be 1"}))
        sys.exit(2)
    cfg = CONFIGS[args.config]
    n = args.n or cfg["n"]
    import torch

    from ablate_amd import fv, synthetic
    from ablate_amd.mesh import BoxMesh

    copy = copy_probe(args.copy_gib) if args.copy_gib > 0 else None
    sprobe = stream_probe(4.0) if args.copy_gib > 0 else None
    mu = 1.8e-5
    cp = 1.4 * 287.0 / 0.4
    k = mu * cp / 0.72
    t_setup = time.perf_counter()
    ordering = args.ordering if args.ordering != "auto" else ("column" if args.strategy == "march" else "brick")
    mesh = BoxMesh([n, n, n], [0, 0, 0], [1, 1, 1], [True, True, True], ordering=ordering)
    solver = fv.FiniteVolumeSolver(mesh, ns_processes(fv, mu, k),
                                   {"petscfv_type": "leastsquares", "petsclimiter_type": "barthjespersen"},
                                   form=cfg["form"])
    solver.set_strategy(args.strategy)
    strategy = solver.strategy()
    U = torch.tensor(synthetic.taylor_green(mesh, cfg["mach"]), dtype=torch.float64, device="cuda")
    R = torch.zeros_like(U)
    info = solver.info()
    setup_s = time.perf_counter() - t_setup
    stream = torch.cuda.current_stream()

    for _ in range(args.warmup):
        solver.compute_rhs(0.0, U, R)
    torch.cuda.synchronize()
    solver.phase_times(reset=True)
    solver.set_profiling(True)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        solver.compute_rhs(0.0, U, R)
    ev1.record(stream)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    solver.set_profiling(False)
    phases = solver.phase_times(reset=True)
    wall_ms = (t1 - t0) * 1e3 / args.steps
    event_ms = ev0.elapsed_time(ev1) / args.steps
    kernel_ms = sum(ms for ms, _ in phases.values()) / args.steps
    cells = info.n_interior
    value = cells / (wall_ms * 1e-3) / 1e6
    alg_bytes = info.alg_bytes_per_cell * cells
    achieved = alg_bytes / (kernel_ms * 1e-3) / 1e9
    finite = bool(torch.isfinite(R).all().item())
    # per-call event timing after the timed region (the measurement plan's median of >= 20 calls): every call is
    # bracketed by its own event pair on the launch stream
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    for a, b in evs:
        a.record(stream)
        solver.compute_rhs(0.0, U, R)
        b.record(stream)
    torch.cuda.synchronize()
    per_call = sorted(a.elapsed_time(b) for a, b in evs)
    median_ms = per_call[len(per_call) // 2] if per_call else None

    out = {
        "metric": "RHS Mcell/s + %HBM peak, FP32-mixed vs FP64 at <=1e-6 rel err; adjoint Mcell/s",
        "value": round(value, 3),
        "unit": "Mcell/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall_ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if cfg["form"] == "fp64" else "f64+f32",
        "data": "synthetic (seeded Taylor-Green state, generated periodic hex mesh)",
        "config": {
            "workload": cfg["workload"],
            "cells": cells,
            "faces": info.n_faces,
            "mesh": f"{n}^3 periodic hex, {ordering}-ordered",
            "form": cfg["form"],
            "strategy": strategy,
            "physics": "NavierStokesTransport(AUSM) + Constant(mu=1.8e-5, Pr=0.72), LSQ + Barth-Jespersen, Mach "
                       f"{cfg['mach']}",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": measured_traffic(cfg, n, strategy),
            "traffic_source": ("rocprofv3 FETCH_SIZE/WRITE_SIZE record in profiles/traffic.json for these device "
                               "sources" if measured_traffic(cfg, n, strategy) is not None
                               else "unmeasured: no counter record of these device sources"),
            # the counter bytes over the same kernel time: how busy the memory system is on the bytes it moves
            "traffic_GB_per_s": (None if measured_traffic(cfg, n, strategy) is None
                                 else round(measured_traffic(cfg, n, strategy) / (kernel_ms * 1e-3) / 1e9, 1)),
            "kernel": "one RHS launch sequence (" + " + ".join(KERNELS[strategy]) + "), event-timed per "
                      f"launch on its stream; algorithmic bytes {info.alg_bytes_per_cell:.1f} B/cell x {cells} cells",
        },
        "phases_ms": {name: round(ms / max(cnt, 1), 4) for name, (ms, cnt) in phases.items() if cnt},
        "event_ms_per_step": round(event_ms, 4),
        "event_ms_median_per_call": None if median_ms is None else round(median_ms, 4),
        "setup_s": round(setup_s, 2),
        "hbm_copy_probe": copy,
        "hbm_stream_probe": sprobe,
        "device_bytes": info.device_bytes,
        "finite": finite,
    }
    progress(f"headline leg done: {value:.1f} Mcell/s")
    if cfg["form"] == "fp64" and not args.no_mixed:
        out["mixed"] = mixed_leg(fv, synthetic, mesh, mu, k, R, U, cfg["mach"], args, wall_ms)
    if cfg["form"] == "fp64" and not args.no_adjoint:
        del solver
        out["adjoint"] = adjoint_leg(fv, synthetic, args)
        progress("adjoint leg done")
    if cfg["form"] == "fp64" and not args.no_tets:
        out["tets"] = tets_leg(fv, synthetic, args)
        progress("tets leg done")
    if cfg["form"] == "fp64" and not args.no_c1:
        out["c1"] = c1_leg(fv, synthetic, args)
        progress("64x64 Euler leg done")
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args.cpu_n, args.cpu_seconds, mu, k, cfg["mach"])
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
