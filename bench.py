"""Throughput of the SIMPLE iteration on one MI355X.

Workload: the 301x701 differentially heated tall cavity (tanh beta = 1.2
mesh, H/W = 2.33, Ra = 1e5, Pr = 0.71, dt = 0.003), all fields resident in
HBM.  A step is one SIMPLE iteration (one time step with single-pass inner
loops): the u, v, p', T line sweeps, the correction, the residual and the
norms.  The same workload also runs on the two FP32 paths, each reported
with its speed-up and its error against FP64:
  - "fp32_mixed": u, v, T, p' stored and swept in FP32; mass residual, p'
    right-hand side, reductions, p and the Nusselt integral in FP64;
  - "fp32_refined": every field in FP64; the sweeps assemble in FP32 and
    solve for FP32 increments against each line's FP64 residual;
  - "fp32_stochastic": the storage of fp32_mixed and the increment sweeps of
    fp32_refined, every new u, v, T stored with stochastic rounding; with it
    the FP32 rounding floor of the u residual and a run to residual 1e-8 that
    finishes on fp32_refined.
Prints one JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--solver pcr|thomas]
  python bench.py --workload batch256   # 256 mixed-size cavities, one launch per phase
  python bench.py --workload adjoint129 # hot-wall Nu gradient through the adjoint
  python bench.py --workload fp32_convergence  # FP64 vs the FP32 paths toward the steady state
  python bench.py --workload barrier --steps 2000  # grid barrier vs kernel boundary (probe)
  python bench.py --workload fortran_host  # the 301x701 case through the Fortran host and its files
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import shutil
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))

HBM_PEAK_GBS = 8000.0

# operands each kernel must read or write once per launch, as (FP64 arrays,
# arrays of the path's real type); every array holds (mi+1)(nj+1) values.  T
# is read by the v sweeps only with buoyancy.  In the FP64 and FP32-refined
# paths both columns are 8-byte; in the FP32-mixed path the second column is
# 4-byte (u, v, T, p', a_u, a_v) and p, b_o, Res_u stay FP64.
ALG_OPERANDS = {
    "sweep_u_y": (1, 4),   # u, u_ant, v, p -> u
    "sweep_u_x": (1, 5),   # u, u_ant, v, p -> u, a_u
    "sweep_v_y": (1, 5),   # v, v_ant, u, p, T -> v
    "sweep_v_x": (1, 7),   # v, v_ant, u, p, T, v at pass start -> v, a_v
    "sweep_p_y": (0, 5),   # u, v, a_u, a_v -> p'
    "sweep_p_x": (1, 6),   # u, v, a_u, a_v, p' -> p', b_o
    "correct": (2, 7),     # p', a_u, a_v, u, v, p -> u, v, p
    "sweep_t_y": (0, 5),   # T, T_ant, u, v -> T
    "sweep_t_x": (0, 6),   # T, T_ant, u, v, T at pass start -> T
    "residual_u": (2, 3),  # u, u_ant, v, p -> Res_u
    "finalize": (0, 0),
}


def alg_bytes(kernel: str, npts: int, precision: str, buoyancy: bool) -> int:
    n64, nr = ALG_OPERANDS[kernel]
    if not buoyancy and kernel in ("sweep_v_y", "sweep_v_x"):
        nr -= 1
    return (8 * n64 + (4 if precision in ("fp32_mixed", "fp32_stochastic") else 8) * nr) * npts


def build_case():
    import cases as K
    return K.nc301x701()


def host_threads() -> int:
    """Host threads the CPU baselines use: OMP_NUM_THREADS when set, else the
    cores this process may run on."""
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        return int(env)
    return len(os.sched_getaffinity(0))


def oracle_timing(case_expr: str, iters_min: int, iters_max: int, budget_s: float, threads: int) -> dict:
    """Seconds per SIMPLE iteration of the C oracle (OpenMP over the loop
    nests) on the case `case_expr` builds, in a child process started with
    OMP_NUM_THREADS = threads and OMP_PROC_BIND = close (the OpenMP runtime
    reads them when it starts; this process has torch's loaded already)."""
    code = (
        "import json, sys, time\n"
        f"sys.path.insert(0, {ROOT!r}); sys.path.insert(0, {os.path.join(ROOT, 'tests', 'golden')!r})\n"
        "import cases as K\n"
        "from ixchel2d_amd import case as C\n"
        "from oracle.pyoracle import Oracle\n"
        f"o = Oracle({case_expr})\n"
        "t = time.perf_counter(); o.run(2); per = (time.perf_counter() - t) / 2\n"
        f"n = int(max({iters_min}, min({iters_max}, {budget_s} / per)))\n"
        "t = time.perf_counter(); o.run(n); per = (time.perf_counter() - t) / n\n"
        "print(json.dumps({'per': per, 'n': n}))\n")
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close")
    r = subprocess.run([sys.executable, "-c", code], env=env, stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                       text=True, check=True)
    return json.loads(r.stdout.strip().splitlines()[-1])


def reference_serial(case, budget_s: float = 15.0):
    """Upstream's own CPU build (oracle/_ref, flang -O2, serial) on a bounded
    number of iterations of the case, or None where it is not built."""
    m = case.mesh
    ref = os.path.join(ROOT, "oracle", "_ref", f"ixchel2d_ref_m{m.mi}_n{m.nj}")
    if not os.path.exists(ref):
        return None
    from ixchel2d_amd import case as C
    work = tempfile.mkdtemp(prefix="ixh_cpu_")

    def run(iters):
        case.params.itermax = iters
        case.params.paq_itera = 10
        C.write_reference_case(case, work, postprocesar=False)
        t = time.perf_counter()
        subprocess.run([ref], cwd=work, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL, check=True)
        return time.perf_counter() - t
    t0 = run(0)
    t1 = run(10)
    per = max((t1 - t0) / 10.0, 1e-6)
    n = int(max(10, min(400, (budget_s / per) // 10 * 10)))
    per = (run(n) - t0) / n
    return {"value": m.mi * m.nj / per / 1e6, "unit": "Mcell-SIMPLE-iters/s", "cores": 1, "kind": "reference",
            "sample": f"upstream serial flang build, {n} iterations of the 301x701 case minus a "
                      f"0-iteration run (setup and file reads)"}


def fortran_host_bench(args):
    """The 301x701 case through the Fortran host (ixchel2d_amd/fortran/
    ixchel2d_gpu: upstream's driver over the C-ABI, reading and writing the
    reference's files).  Two runs in packages of 1000 iterations, of 1000
    and of 1000 + steps iterations (steps rounded up to whole packages, at
    least 4000), each the fastest of three; the difference times the steps
    with their packages' outputs (fort.102, the Nusselt record, out_*.dat,
    the .bin dumps and the VTK sample) and without process start, file reads
    and set-up.  The same pair again without the
    package outputs (postprocesar = .false.: fort.102 and the Nusselt record
    only) shows what the outputs cost.  Upstream's serial build on the same
    case beside it."""
    from ixchel2d_amd import case as C
    exe = os.path.join(ROOT, "ixchel2d_amd", "fortran", "ixchel2d_gpu")
    if not os.path.exists(exe):
        raise SystemExit("the Fortran host is not built (python -c 'import __graft_entry__ as g; g.build()')")
    paq = 1000
    steps = max(4 * paq, -(-args.steps // paq) * paq)
    case = build_case()
    env = dict(os.environ, IXCHEL2D_SOLVER=args.solver, IXCHEL2D_PRECISION=args.precision)

    def run(iters, outputs=True):
        work = tempfile.mkdtemp(prefix="ixh_fhost_")
        try:
            case.params.itermax = iters
            case.params.paq_itera = paq
            C.write_reference_case(case, work, postprocesar=outputs)
            t = time.perf_counter()
            subprocess.run([exe], cwd=work, env=env, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL,
                           check=True, timeout=600)
            return time.perf_counter() - t
        finally:
            shutil.rmtree(work, ignore_errors=True)
    # each process time is the least of three runs: process start and file
    # reads vary by tens of ms from run to run
    t_base = min(run(paq) for _ in range(3))
    t_full = min(run(paq + steps) for _ in range(3))
    per = (t_full - t_base) / steps
    t_base0 = min(run(paq, outputs=False) for _ in range(3))
    t_full0 = min(run(paq + steps, outputs=False) for _ in range(3))
    per0 = (t_full0 - t_base0) / steps
    m = case.mesh
    cpu = None if args.no_cpu else reference_serial(build_case())
    out = {
        "metric": "Mcell-SIMPLE-iters/s @301x701 through the Fortran host (reference files in and out)",
        "value": round(m.mi * m.nj / per / 1e6, 2), "unit": "Mcell-SIMPLE-iters/s",
        "n_gpus": 1, "steps": steps, "warmup": paq, "ms_per_step": round(per * 1e3, 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32+f64", "data": "synthetic",
        "config": {"workload": "natural-convection tall cavity 301x701 (tanh beta=1.2, H/W=2.33, Ra=1e5, "
                               "Pr=0.71, dt=0.003) through ixchel2d_gpu, packages of 1000 iterations with "
                               "their outputs", "solver": args.solver, "path": args.precision,
                   "process_seconds": {f"{paq} iterations": round(t_base, 3),
                                       f"{paq + steps} iterations": round(t_full, 3)}},
        "without_package_outputs": {"value": round(m.mi * m.nj / per0 / 1e6, 2), "ms_per_step": round(per0 * 1e3, 5),
                                    "process_seconds": {f"{paq} iterations": round(t_base0, 3),
                                                        f"{paq + steps} iterations": round(t_full0, 3)}},
        "package_output_seconds": round((per - per0) * paq, 3),
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


def cpu_baseline(case, budget_s: float = 15.0):
    """The multicore CPU baseline: the C oracle (the FP64 restatement of the
    same loop nests, OpenMP) on every host thread, on a bounded number of
    iterations of the same case; upstream's serial build beside it.
    Upstream's multicore build cannot be made: its parallelism is OpenACC,
    which flang rejects, there is no gfortran or nvfortran here, and the
    source has no OpenMP."""
    m = case.mesh
    threads = host_threads()
    t = oracle_timing("K.nc301x701()", 5, 2000, budget_s, threads)
    out = {"value": m.mi * m.nj / t["per"] / 1e6, "unit": "Mcell-SIMPLE-iters/s", "cores": threads,
           "kind": "port",
           "sample": f"C oracle (FP64 restatement, OpenMP, OMP_PROC_BIND=close), {t['n']} iterations of the "
                     "301x701 case on every host thread",
           "note": "upstream's multicore build does not exist here: its parallelism is OpenACC (flang rejects "
                   "it, no gfortran/nvfortran), and its source has no OpenMP; the serial upstream build is "
                   "timed beside"}
    out["reference_serial"] = reference_serial(case)
    return out


def batch_traffic(args):
    """HBM bytes of one step of the default 256-cavity FP64 batch, from the
    FETCH_SIZE / WRITE_SIZE counter runs summarised in profiles/ (None for
    other batches and precisions, which were not counted)."""
    path = os.path.join(ROOT, "profiles", f"traffic_batch256_{args.precision}.json")
    if args.batch != 256 or args.solver != "pcr" or not os.path.exists(path):
        return None
    with open(path) as fh:
        return json.load(fh).get("hbm_bytes_per_step")


def batch_bench(args):
    """Aggregate throughput of the mixed-size batch (256 cavities of 65^2,
    129^2, 257^2 and 129x301 cells, lid-driven and natural convection, 2 of
    16 with a cylinder), schedule from the autotuner."""
    import torch
    from ixchel2d_amd import case as C
    from ixchel2d_amd.simple import Batch, Cavity

    cases = C.batch_cases(args.batch, seed=0)
    cav = [Cavity(c, solver=args.solver, precision=args.precision) for c in cases]
    b = Batch(cav)
    trial = 3
    sched, table = b.autotune(trial_steps=trial)
    b.run(args.warmup, cap=1)
    torch.cuda.synchronize()
    warm_steps = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < args.warm_seconds:
        b.run(20, cap=1)
        torch.cuda.synchronize()
        warm_steps += 20
    warm_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    b.run(args.steps, cap=1)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    cells = sum(c.mesh.mi * c.mesh.nj for c in cases)
    ms_per_step = dt * 1e3 / args.steps
    import numpy as np
    nonfinite = [c.case.name for c in cav if not all(np.all(np.isfinite(c.get(f))) for f in "uvtp")]
    # algorithmic bytes of one step: the per-kernel array counts over every
    # cavity; the v sweeps read T only with buoyancy; sources (cylinders) add
    # Sc, Sp to the u, v, T sweeps and to the residual
    iter_bytes = 0
    for c in cases:
        npts = (c.mesh.mi + 1) * (c.mesh.nj + 1)
        buoy = c.params.derived()["Riy"] != 0.0
        iter_bytes += sum(alg_bytes(k, npts, args.precision, buoy) for k in ALG_OPERANDS)
        if c.sources:
            iter_bytes += 2 * 7 * 8 * npts
    gbs = iter_bytes / (ms_per_step * 1e-3) / 1e9
    total_steps = len(table) * (trial + 4) + args.warmup + warm_steps + args.steps
    nu_err = None
    if args.precision != "fp64":
        # every cavity's hot-wall Nusselt number against the same batch run
        # on the FP64 path for the same steps
        import numpy as np
        nu32 = [c.nusselt_x()[0] for c in cav]
        b.close()
        for c in cav:
            c.close()
        ref = [Cavity(c, solver=args.solver, precision="fp64") for c in cases]
        rb = Batch(ref, order="xcd")
        rb.run(total_steps, cap=1)
        rb.close()
        errs = []
        for k, r in enumerate(ref):
            nu64 = r.nusselt_x()[0]
            if abs(nu64) > 1e-12:
                errs.append(abs(nu32[k] - nu64) / abs(nu64))
            r.close()
        e = np.array(errs)
        nu_err = {"cavities_with_heat_flow": len(errs), "steps": total_steps, "max": float(e.max()),
                  "median": float(np.median(e)), "p90": float(np.quantile(e, 0.9))}
    cpu = None
    if not args.no_cpu:
        from oracle.pyoracle import Oracle
        sample = cases[:16]
        threads = host_threads()
        t = time.perf_counter()
        for c in sample:
            Oracle(c).run(10)
        per = time.perf_counter() - t
        scells = sum(c.mesh.mi * c.mesh.nj for c in sample)
        cpu = {"value": scells * 10 / per / 1e6, "unit": "Mcell-SIMPLE-iters/s", "cores": threads, "kind": "port",
               "sample": "C oracle (OpenMP), 10 iterations of each of the first 16 cavities, one after another"}
    out = {
        "metric": "aggregate Mcell-SIMPLE-iters/s, 256 mixed-size cavities",
        "value": round(cells * args.steps / dt / 1e6, 2),
        "unit": "Mcell-SIMPLE-iters/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64" if args.precision == "fp64" else "f32+f64", "data": "synthetic",
        "config": {"workload": f"{args.batch} cavities: sizes 65^2/129^2/257^2/129x301 cells, half lid-driven "
                               "Re 100..3200, half natural convection Ra 1e3..1e6, 2 of 16 with a cylinder",
                   "cavities": args.batch, "cells": cells, "solver": args.solver, "path": args.precision,
                   "schedule": sched,
                   "autotune_ms_per_step": {
                       f"{k[0]}{'+graph' if k[1] else ''}/{k[2]}t{f'/{k[3]}lines' if k[3] > 1 else ''}"
                       f"{'/res-with-T' if k[4] else ''}": round(v, 4)
                       for k, v in table.items()}},
        "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": batch_traffic(args), "kernel": "whole step",
                     "iteration_alg_bytes": iter_bytes},
        "warm_untimed": {"steps": warm_steps, "seconds": round(warm_s, 3)},
        "nu_rel_err_vs_fp64": nu_err,
        "nonfinite_cavities": nonfinite,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


def adjoint_bench(args):
    """Gradient of the mean hot-wall Nusselt number with respect to 64
    piecewise-linear hot-wall temperature controls (base 1 +- U(0.1)) after
    K = 400 SIMPLE iterations from a converged start, 129^2-cell natural
    convection (Ra = 1e5, Pr = 0.71, uniform mesh), checkpoints every 20
    iterations, through the adjoint of the whole step: gradient evaluations
    per second, device memory, and the gradient against central differences
    (h = 1e-6), with buoyancy and without."""
    import numpy as np
    import torch
    from ixchel2d_amd import adjoint as A
    from ixchel2d_amd import case as C
    from ixchel2d_amd.simple import Cavity

    case = C.natural_convection_case(130, 130, height=1.0, Ra=1e5, iters=400, dt=args.adj_dt, seed=0,
                                     mesh_kind="unif")
    warm = Cavity(case, solver="pcr")
    hw = warm.run(args.adj_warm, cap=1)
    hw2 = warm.run(1)
    start = {f: warm.get(f) for f in ("u", "v", "t", "p")}
    warm.close()
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    prob = A.HotWallProblem(case, ncontrols=64, steps=400, every=20, start=start)
    rng = np.random.default_rng(0)
    c = 1.0 + rng.uniform(-0.1, 0.1, 64)
    prob.forward(c)
    prob.gradient()
    torch.cuda.synchronize()
    reps = args.adj_reps
    t0 = time.perf_counter()
    for _ in range(reps):
        nu = prob.forward(c)
        g = prob.gradient()
    torch.cuda.synchronize()
    per = (time.perf_counter() - t0) / reps
    used = free0 - torch.cuda.mem_get_info()[0]

    def fd(prob, h=1e-6):
        out = np.zeros(64)
        for i in range(64):
            e = np.zeros(64)
            e[i] = h
            out[i] = (prob.forward(c + e) - prob.forward(c - e)) / (2 * h)
        return out

    t1 = time.perf_counter()
    fd_coupled = fd(prob)
    fd_time = time.perf_counter() - t1
    prob.forward(c)
    g = prob.gradient()
    rel_coupled = float(np.max(np.abs(g - fd_coupled)) / np.max(np.abs(fd_coupled)))
    prob.coupled = False
    g_frozen = prob.gradient()
    rel_frozen = float(np.max(np.abs(g_frozen - fd_coupled)) / np.max(np.abs(fd_coupled)))
    t3 = time.perf_counter()
    for _ in range(reps):
        prob.forward(c)
        prob.gradient()
    torch.cuda.synchronize()
    per_frozen = (time.perf_counter() - t3) / reps
    prob.coupled = True
    prob.cav.set_params(ri_y=0.0)
    prob.forward(c)
    g0 = prob.gradient()
    fd_passive = fd(prob)
    rel_passive = float(np.max(np.abs(g0 - fd_passive)) / np.max(np.abs(fd_passive)))
    cpu = None
    if not args.no_cpu:
        from oracle.pyoracle import Oracle
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))

        def J(cc):
            o = Oracle(case)
            o.set_bc("t", "a", prob.types_a, prob.wall_values(cc))
            for f, arr in start.items():
                o.set(f, arr)
                if f in ("u", "v", "t"):
                    o.set(f + "_ant", arr)
            o.run(400)
            return o.nusselt_x()[0]
        nsample = 4
        t2 = time.perf_counter()
        ref = []
        for i in range(nsample):
            e = np.zeros(64)
            e[i] = 1e-6
            ref.append((J(c + e) - J(c - e)) / 2e-6)
        tc = time.perf_counter() - t2
        cpu = {"value": round(1.0 / (tc / nsample * 64), 5), "unit": "gradient evaluations/s", "cores": threads,
               "kind": "port",
               "sample": f"C oracle (OpenMP) central differences of the coupled run for {nsample} of the 64 "
                         "controls (2 x 400 iterations each), scaled to 64 controls",
               "fd_reference_vs_gpu_fd_rel": float(np.max(np.abs(np.array(ref) - fd_coupled[:nsample]))
                                                   / np.max(np.abs(fd_coupled[:nsample])))}
    out = {
        "metric": "gradient evaluations/s, d(hot-wall Nu)/d(64 wall-temperature controls), K=400",
        "value": round(1.0 / per, 3), "unit": "gradient evaluations/s", "n_gpus": 1, "steps": reps,
        "warmup": 1, "ms_per_step": round(per * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "natural convection 129^2 cells uniform, Ra=1e5, Pr=0.71, K=400 iterations "
                               "from a converged start, checkpoints every 20",
                   "dt": args.adj_dt, "warm_iterations": args.adj_warm,
                   "start_residuo": float(hw2["residuo"][-1]), "start_maxbo": float(hw2["maxbo"][-1])},
        "nu": nu,
        "device_memory_MB": round(used / 2**20, 1),
        "gradient_check": {
            "adjoint_rel_err_vs_central_fd": rel_coupled,
            "without_buoyancy_rel_err_vs_central_fd": rel_passive,
            "energy_only_adjoint_rel_err_vs_central_fd": rel_frozen,
            "energy_only_gradient_evaluations_per_s": round(1.0 / per_frozen, 3),
            "fd_seconds_64_controls": round(fd_time, 2),
            "note": "adjoint of the whole step (momentum, p', correction, energy); the energy-only adjoint "
                    "holds the velocities fixed"},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


LAUNCH_PARTS = {
    "multi_v_y+residual": ("sweep_v_y", "residual_u"),
    "multi_v_x+fold": ("sweep_v_x", "finalize"),
    "step_sweep_p_y": ("sweep_p_y",),
    "step_p_x+correct": ("sweep_p_x", "correct"),
    "multi_t_y+u_y": ("sweep_t_y", "sweep_u_y"),
    "multi_t_x+u_x": ("sweep_t_x", "sweep_u_x"),
}


def convergence_study(args):
    """FP64 and both FP32 paths on the 301x701 case side by side toward the
    steady state: residual, maxbo, hot-wall Nu and its relative difference
    from FP64, and the max-norm differences of T and u, every --conv-chunk
    iterations."""
    import numpy as np
    import torch
    from ixchel2d_amd.simple import Cavity
    case = build_case()
    paths = ("fp64", "fp32_mixed", "fp32_refined", "fp32_stochastic")
    cav = {p: Cavity(case, solver="pcr", precision=p) for p in paths}
    rows = []
    total = 0
    for _ in range(args.conv_chunks):
        hist = {}
        for p in paths:
            cav[p].run(args.conv_chunk - 1, cap=1)
            hist[p] = cav[p].run(1)
        total += args.conv_chunk
        nu = {p: cav[p].nusselt_x()[0] for p in paths}
        t64, u64 = cav["fp64"].get("t"), cav["fp64"].get("u")
        row = {"iterations": total}
        for p in paths:
            row[f"residuo_{p}"] = float(hist[p]["residuo"][-1])
            row[f"maxbo_{p}"] = float(hist[p]["maxbo"][-1])
            row[f"nu_{p}"] = nu[p]
        for p in paths[1:]:
            row[f"nu_rel_err_{p}"] = abs(nu[p] - nu["fp64"]) / abs(nu["fp64"])
            row[f"t_linf_{p}"] = float(np.max(np.abs(cav[p].get("t") - t64)))
            row[f"u_linf_rel_{p}"] = float(np.max(np.abs(cav[p].get("u") - u64)) / np.max(np.abs(u64)))
        rows.append(row)
        print(json.dumps(row), file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    print(json.dumps({"workload": "natural-convection tall cavity 301x701, FP64 and the FP32 paths "
                                  "toward the steady state", "chunk": args.conv_chunk, "rows": rows}))


def time_path(case, solver, precision, steps, warmup, warm_s=1.0):
    """Whole-step throughput of one cavity and the average kernel duration of
    each of the step's 6 launches over the same timed steps."""
    import torch
    from ixchel2d_amd.simple import Cavity
    m = case.mesh
    cells = m.mi * m.nj
    npts = (m.mi + 1) * (m.nj + 1)
    gpu = Cavity(case, solver=solver, precision=precision)
    gpu.run(warmup, cap=1)
    torch.cuda.synchronize()
    # untimed steps until warm_s seconds have passed, so that the timed
    # steps run at steady clocks however few they are
    warm_steps = 0
    tw = time.perf_counter()
    while time.perf_counter() - tw < warm_s:
        gpu.run(200, cap=1)
        torch.cuda.synchronize()
        warm_steps += 200
    warm_took = time.perf_counter() - tw
    t0 = time.perf_counter()
    gpu.run(steps)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ms_per_step = dt * 1e3 / steps
    # the same number of steps again with every launch recording its
    # kernel's start and end (HIP events on the library's stream, read back
    # from a ring).  The events slow the host-side launches, so this run
    # gives the kernel durations and the timed run above the throughput.
    gpu.set_profiling(True)
    t1 = time.perf_counter()
    gpu.run(steps)
    torch.cuda.synchronize()
    ms_events = (time.perf_counter() - t1) * 1e3 / steps
    kt = gpu.kernel_times()
    gpu.set_profiling(False)
    buoy = case.params.derived()["Riy"] != 0.0
    launches = {}
    for name, (ms, n) in kt.items():
        if name not in LAUNCH_PARTS or n == 0:
            continue
        us = ms * 1e3 / n
        b = sum(alg_bytes(k, npts, precision, buoy) for k in LAUNCH_PARTS[name])
        launches[name] = {"avg_us": round(us, 3), "launches": n, "alg_bytes": b,
                          "GB/s": round(b / (us * 1e-6) / 1e9, 1)}
    dom = max(launches, key=lambda k: launches[k]["avg_us"])
    dk = launches[dom]
    iter_bytes = sum(alg_bytes(k, npts, precision, buoy) for k in ALG_OPERANDS)
    it_gbs = iter_bytes / (ms_per_step * 1e-3) / 1e9
    # HBM bytes per launch from the committed counter profile of this
    # workload (FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0, scaled on a 512 MiB copy)
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_nc301x701_fused.json")
    if os.path.exists(tpath) and (m.mi, m.nj) == (301, 701):
        with open(tpath) as fh:
            traffic = json.load(fh).get(precision, {}).get(dom, {}).get("hbm_bytes")
    roofline = {"bound": "hbm", "achieved": dk["GB/s"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(dk["GB/s"] / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": dom,
                "alg_bytes_per_launch": dk["alg_bytes"], "avg_launch_us": dk["avg_us"],
                "iteration_alg_bytes": iter_bytes, "iteration_GB/s": round(it_gbs, 1),
                "iteration_frac": round(it_gbs / HBM_PEAK_GBS, 4),
                "sum_launch_us": round(sum(v["avg_us"] for v in launches.values()), 2)}
    return {"value": round(cells * steps / dt / 1e6, 2), "ms_per_step": round(ms_per_step, 5),
            "ms_per_step_with_events": round(ms_events, 5), "roofline": roofline, "launches": launches,
            "warm_untimed": {"steps": warm_steps, "seconds": round(warm_took, 3)}}


def accuracy(case, solver, steps, precision):
    """One FP32 path against FP64 from the same start after `steps`
    iterations (a transient): relative error of the hot-wall mean Nusselt
    number and max-norm errors of u, v, T (relative to the FP64 field's max)."""
    from ixchel2d_amd.simple import Cavity
    import numpy as np
    res = {}
    for prec in ("fp64", precision):
        g = Cavity(case, solver=solver, precision=prec)
        h = g.run(steps, cap=1)
        h = g.run(1)
        res[prec] = ({f: g.get(f) for f in "uvt"}, g.nusselt_x()[0], float(h["residuo"][-1]))
        g.close()
    f64, nu64, r64 = res["fp64"]
    f32, nu32, r32 = res[precision]
    out = {"iterations": steps + 1, "nu_hot_fp64": nu64, f"nu_hot_{precision}": nu32,
           "nu_rel_err": abs(nu32 - nu64) / abs(nu64), "residuo_fp64": r64, f"residuo_{precision}": r32}
    for f in "uvt":
        out[f"linf_rel_err_{f}"] = float(np.max(np.abs(f32[f] - f64[f])) / np.max(np.abs(f64[f])))
    return out


def run_to_residual(case, solver, precision, tol, chunk, max_iters):
    """Iterate until the FP64 u residual norm (residuo_u of the path's own
    fields, computed in FP64 on every path) is at most tol, checked every
    `chunk` iterations, or until max_iters.  Returns the cavity, the
    iterations, the last residual and maxbo and the wall time."""
    import torch
    from ixchel2d_amd.simple import Cavity
    g = Cavity(case, solver=solver, precision=precision)
    it = 0
    t = time.perf_counter()
    while True:
        g.run(chunk - 1, cap=1)
        h = g.run(1)
        it += chunk
        r, mb = float(h["residuo"][-1]), float(h["maxbo"][-1])
        if r <= tol or it >= max_iters:
            break
    torch.cuda.synchronize()
    return g, it, r, mb, time.perf_counter() - t


def accuracy_converged(ref, case, solver, precision, tol, chunk, max_iters):
    """An FP32 path run to the FP64-measured residual tol, against the FP64
    path run to the same tol (`ref`: cavity, iterations, residual, maxbo,
    seconds): hot-wall Nusselt number, max-norm field errors."""
    import numpy as np
    g64, it64, r64, mb64, s64 = ref
    g, it, r, mb, sec = run_to_residual(case, solver, precision, tol, chunk, max_iters)
    nu64, nu = g64.nusselt_x()[0], g.nusselt_x()[0]
    out = {"tol": tol, "iterations": it, "residuo": r, "maxbo": mb, "reached": r <= tol, "seconds": round(sec, 3),
           "fp64_iterations": it64, "fp64_residuo": r64, "fp64_seconds": round(s64, 3),
           "nu_hot_fp64": nu64, f"nu_hot_{precision}": nu, "nu_rel_err": abs(nu - nu64) / abs(nu64)}
    for f in "uvt":
        a, b = g.get(f), g64.get(f)
        out[f"linf_rel_err_{f}"] = float(np.max(np.abs(a - b)) / np.max(np.abs(b)))
    g.close()
    return out


STATE_FIELDS = ("u", "v", "t", "p", "u_ant", "v_ant", "t_ant")


def rounding_floor(ref, case):
    """What FP32 storage of u costs the FP64-measured residual: the steady u
    residual (u_ant = u) of the FP64 path's converged state, and of the same
    state with u, v, T rounded to FP32, to nearest and stochastically (u_ant
    then rounded on its own, as two steps of a stochastic iteration would
    leave it).  No FP32-stored u can go below the rounded figures."""
    import numpy as np
    from ixchel2d_amd.simple import Cavity
    g64 = ref[0]
    F = {n: g64.get(n) for n in STATE_FIELDS}
    rng = np.random.default_rng(0)

    def rn(a):
        return a.astype(np.float32).astype(np.float64)

    def sr(a):
        f = a.astype(np.float32)
        fd = f.astype(np.float64)
        lo = np.where(fd > a, np.nextafter(f, np.float32(-np.inf)), f).astype(np.float64)
        hi = np.where(fd < a, np.nextafter(f, np.float32(np.inf)), f).astype(np.float64)
        w = np.where(hi > lo, (a - lo) / np.where(hi > lo, hi - lo, 1.0), 0.0)
        return np.where(rng.random(a.shape) < w, hi, lo)

    def resid(fields):
        c = Cavity(case, solver="pcr", precision="fp64", upload=False)
        for n, a in fields.items():
            c.set(n, a)
        r = c.residual_u()
        c.close()
        return r

    exact = dict(F, u_ant=F["u"], v_ant=F["v"], t_ant=F["t"])
    out = {"fp64_iterations": ref[1], "residuo_fp64_state": resid(exact)}
    for name, fn in (("nearest", rn), ("stochastic", sr)):
        m = {k: fn(F[k]) if k in ("u", "v", "t") else F[k] for k in ("u", "v", "t", "p")}
        m.update(u_ant=fn(F["u"]) if name == "stochastic" else m["u"], v_ant=m["v"], t_ant=m["t"])
        out[f"residuo_rounded_{name}"] = resid(m)
    return out


def hybrid_to_residual(ref, case, tol, chunk, max_iters):
    """The FP32-stochastic path until its residual is within 3 tol (or has
    not fallen 10% below its best for 3 chunks: FP32 storage of u keeps it
    above a floor), then the FP32-refined path from its fields until the
    FP64-measured residual is tol: FP32 arithmetic throughout, FP64 storage
    only for the last stretch.  Against the FP64 path run to the same tol."""
    import numpy as np
    import torch
    from ixchel2d_amd.simple import Cavity
    g64, it64, r64, mb64, s64 = ref
    t0 = time.perf_counter()
    g = Cavity(case, solver="pcr", precision="fp32_stochastic")
    it1, best, flat, r = 0, np.inf, 0, np.inf
    while it1 < max_iters:
        g.run(chunk - 1, cap=1)
        r = float(g.run(1)["residuo"][-1])
        it1 += chunk
        flat = 0 if r < 0.9 * best else flat + 1
        best = min(best, r)
        if r <= 3.0 * tol or flat >= 3:
            break
    r1 = r
    it2 = 0
    if r > tol:
        h = Cavity(case, solver="pcr", precision="fp32_refined", upload=False)
        for n in STATE_FIELDS:
            h.set(n, g.get(n))
        g.close()
        g = h
        while it1 + it2 < max_iters:
            g.run(chunk - 1, cap=1)
            r = float(g.run(1)["residuo"][-1])
            it2 += chunk
            if r <= tol:
                break
    torch.cuda.synchronize()
    sec = time.perf_counter() - t0
    nu64, nu = g64.nusselt_x()[0], g.nusselt_x()[0]
    out = {"tol": tol, "iterations_fp32_stochastic": it1, "residuo_at_switch": r1,
           "iterations_fp32_refined": it2, "residuo": r, "reached": r <= tol, "seconds": round(sec, 3),
           "fp64_iterations": it64, "fp64_seconds": round(s64, 3), "time_speedup_vs_fp64": round(s64 / sec, 3),
           "nu_hot_fp64": nu64, "nu_hot": nu, "nu_rel_err": abs(nu - nu64) / abs(nu64)}
    for f in "uvt":
        a, b = g.get(f), g64.get(f)
        out[f"linf_rel_err_{f}"] = float(np.max(np.abs(a - b)) / np.max(np.abs(b)))
    g.close()
    return out


def barrier_probe(args):
    """Cost of a grid-wide barrier in one cooperative launch against a kernel
    boundary, at the grid shapes of a persistent fused step
    (ixchel2d_amd/csrc/probe_barrier.hip); prints the probe's JSON line."""
    exe = os.path.join(ROOT, "ixchel2d_amd", "probe_barrier")
    if not os.path.exists(exe):
        raise SystemExit(f"{exe} is not built; run __graft_entry__.build()")
    for per in (1, 8):
        r = subprocess.run([exe, str(args.steps), str(per)], stdout=subprocess.PIPE, text=True, check=True,
                           timeout=300)
        print(r.stdout.strip())


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--solver", default="pcr", choices=["pcr", "thomas"])
    ap.add_argument("--workload", default="nc301x701",
                    choices=["nc301x701", "batch256", "adjoint129", "fp32_convergence", "barrier", "fortran_host"])
    ap.add_argument("--conv-chunk", type=int, default=30000, help="iterations per row of the convergence study")
    ap.add_argument("--conv-chunks", type=int, default=12, help="rows of the convergence study")
    ap.add_argument("--adj-dt", type=float, default=0.01, help="time step of the adjoint workload")
    ap.add_argument("--adj-warm", type=int, default=20000, help="iterations to the converged start")
    ap.add_argument("--adj-reps", type=int, default=5, help="timed gradient evaluations")
    ap.add_argument("--batch", type=int, default=256, help="cavities of the batch workload")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32_mixed", "fp32_refined", "fp32_stochastic"],
                    help="path of the batch workload")
    ap.add_argument("--fp64-only", action="store_true", help="skip the FP32 lines of the 301x701 workload")
    ap.add_argument("--accuracy-steps", type=int, default=5000,
                    help="iterations of the transient FP32 vs FP64 accuracy comparisons")
    ap.add_argument("--converge-tol", type=float, default=1e-8,
                    help="FP64-measured residual the FP32 paths and FP64 are run to for their accuracy "
                         "(0: skip)")
    ap.add_argument("--converge-chunk", type=int, default=10000, help="iterations between residual checks")
    ap.add_argument("--warm-seconds", type=float, default=1.0,
                    help="untimed steps for this long after the warm-up steps, before the timed steps")
    ap.add_argument("--calibrate-copy-mb", type=int, default=0,
                    help="first run one device-to-device copy of this many MB (reference kernel for "
                         "the FETCH_SIZE / WRITE_SIZE counter scale)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(a batch would shard across GPUs by cavity; this build runs on one GPU)", file=sys.stderr)
        sys.exit(2)

    if args.workload == "fortran_host":
        fortran_host_bench(args)
        return
    import torch

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    if args.workload == "batch256":
        batch_bench(args)
        return
    if args.workload == "adjoint129":
        adjoint_bench(args)
        return
    if args.workload == "fp32_convergence":
        convergence_study(args)
        return
    if args.workload == "barrier":
        barrier_probe(args)
        return
    if args.calibrate_copy_mb:
        n = args.calibrate_copy_mb * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    case = build_case()
    m = case.mesh
    cells = m.mi * m.nj
    r64 = time_path(case, args.solver, "fp64", args.steps, args.warmup, args.warm_seconds)
    reduced = {"fp32_mixed": None, "fp32_refined": None, "fp32_stochastic": None}
    if args.solver == "pcr" and not args.fp64_only:
        ref = None
        if args.converge_tol > 0:
            ref = run_to_residual(build_case(), "pcr", "fp64", args.converge_tol, args.converge_chunk, 400000)
        for prec in reduced:
            r = time_path(build_case(), "pcr", prec, args.steps, args.warmup, args.warm_seconds)
            r["speedup_vs_fp64"] = round(r["value"] / r64["value"], 3)
            r["accuracy"] = accuracy(build_case(), "pcr", args.accuracy_steps, prec)
            if ref is not None:
                r["accuracy_converged"] = accuracy_converged(ref, build_case(), "pcr", prec, args.converge_tol,
                                                             args.converge_chunk, int(1.5 * ref[1]))
            reduced[prec] = r
        if ref is not None:
            st = reduced["fp32_stochastic"]
            st["fp32_rounding_floor"] = rounding_floor(ref, build_case())
            st["stochastic_then_refined"] = hybrid_to_residual(ref, build_case(), args.converge_tol,
                                                               max(1, args.converge_chunk // 2), int(1.5 * ref[1]))
            ref[0].close()
    cpu = None if args.no_cpu else cpu_baseline(build_case())
    out = {
        "metric": "Mcell-SIMPLE-iters/s @301x701",
        "value": r64["value"],
        "unit": "Mcell-SIMPLE-iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r64["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "natural-convection tall cavity 301x701 (tanh beta=1.2, H/W=2.33, Ra=1e5, "
                               "Pr=0.71, dt=0.003), FP64 plain path, one SIMPLE iteration per step",
                   "mesh": [m.mi, m.nj], "cells": cells, "solver": args.solver, "path": "fp64"},
        "roofline": r64["roofline"],
        "warm_untimed": r64["warm_untimed"],
        "cpu_baseline": cpu,
        "launches": r64["launches"],
        "fp32_mixed": reduced["fp32_mixed"],
        "fp32_refined": reduced["fp32_refined"],
        "fp32_stochastic": reduced["fp32_stochastic"],
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
