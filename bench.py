#!/usr/bin/env python
"""Throughput of the OpenSBLI RHS + RK3 time advance on one MI355X.

Workload: 3-D Taylor-Green vortex, 256^3 periodic, FP64, Re = 1600, M = 0.1,
WENO5-Z characteristic LLF flux (Roe average) with 4th-order central viscous
terms and the 3-register RK3 (the TGV script's RungeKutta(3)).  A step is one
RK3 iteration (three stages).  The state is resident in HBM before timing.

Prints one JSON line:
  value = Mcell-updates/s per RK stage = cells * stages * steps / t / 1e6
  roofline = the dominant kernel against its binding roofline (FP64 issue
             for the characteristic sweeps): its FP64 operations per launch
             (counted from the kernel formulas) / its mean launch time (HIP
             events on the library's stream) vs 78.6 TFLOP/s, or its
             algorithmic bytes per launch vs 8 TB/s for an HBM-bound kernel;
             roofline_fp64 and roofline_hbm give both views
  ensemble = the batched-ensemble speed-up (512 uneven 1-D WENO5-Z shock
             tubes: one persistent launch vs per-instance launches)
  flat_plate = the 480x256x128 flat plate on the tanh (factor 5) wall-normal
             mesh, Mcell-updates/s per RK stage
  adjoint = backward/forward time of the checkpointed 512^2 Kelvin-Helmholtz run
  cpu_baseline = the oracle (C restatement of the generated OPS kernel
             sequence) on the host's cores, one warm-up step then the median
             of three timed steps of the same workload.
"""
import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# measured HBM bytes per launch of each kernel family (rocprofv3 FETCH_SIZE and
# WRITE_SIZE passes over this benchmark, scaled by a plain copy's readings);
# the commands and raw figures are in profiles/hbm_traffic_<scheme>_tgv256.md
TRAFFIC_FILES = [os.path.join(ROOT, "profiles", f"hbm_traffic_{s}_tgv256.json") for s in ("weno_z", "central_skew")]
FP64_PEAK_TFS = 78.6


def _log(msg):
    print(msg, file=sys.stderr, flush=True)


def measured_traffic(family, n, conv):
    """HBM bytes per launch of a kernel family from the committed measurement,
    when it covers this workload; None otherwise."""
    for path in TRAFFIC_FILES:
        try:
            with open(path) as f:
                t = json.load(f)
        except (OSError, ValueError):
            continue
        if t.get("grid") == n and t.get("scheme") == conv:
            fam = t.get("families", {}).get(family)
            return None if fam is None else fam.get("bytes_per_launch")
    return None


def host_cpu():
    """Model, sockets, cores and the cores this process may use (lscpu)."""
    info = {"logical_cpus": os.cpu_count()}
    try:
        info["usable_cpus"] = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        pass
    try:
        txt = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
    except (OSError, subprocess.SubprocessError):
        return info
    keys = {"Model name": "model", "Socket(s)": "sockets", "Core(s) per socket": "cores_per_socket",
            "Thread(s) per core": "threads_per_core"}
    for line in txt.splitlines():
        k, _, v = line.partition(":")
        if k.strip() in keys:
            info[keys[k.strip()]] = v.strip()
    return info


def _median_timed(solver, steps, reps=3):
    """One untimed warm-up run of `steps` (first touch of the work arrays),
    then the median wall time of `reps` timed runs."""
    solver.run(steps)
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        solver.run(steps)
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts)), ts


def cpu_baseline(n, conv, steps=1):
    """Time the oracle on this host, the same problem: a warm-up of `steps`
    RK3 steps, then the median of three timed runs of `steps` steps."""
    from opensbli_amd.problems import taylor_green
    from oracle.ref import RefSolver
    threads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
    problem, q0 = taylor_green(n, conv=conv)
    s = RefSolver(problem, variant="fma")
    s.set_state(q0)
    dt, ts = _median_timed(s, steps)
    return {
        "value": n ** 3 * 3 * steps / dt / 1e6,
        "unit": "Mcell-updates/s per RK stage",
        "cores": threads,
        "kind": "port",
        "sample": f"{steps} RK3 step(s) of the same {n}^3 TGV {conv} problem, median of 3 after 1 warm-up "
                  f"({', '.join('%.2f' % t for t in ts)} s; OMP_NUM_THREADS={threads}; oracle built with -O3 "
                  f"-march=x86-64-v3 -ffp-contract=fast)",
        "host": host_cpu(),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--conv", default="weno_z", choices=["weno_z", "weno_js", "teno6", "teno5", "central_skew"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=1)
    ap.add_argument("--rk", default="rk3", choices=["rk3", "ls3", "ls3_ssp", "ls4"],
                    help="tgv: RungeKutta(3) (3-register) or a low-storage scheme")
    ap.add_argument("--workload", default="tgv", choices=["tgv", "ensemble", "flat_plate", "adjoint",
                                                           "curvilinear", "cpu"])
    ap.add_argument("--no-adjoint", action="store_true", help="tgv: skip the adjoint/forward ratio measurement")
    ap.add_argument("--no-sub", action="store_true", help="tgv: skip the ensemble and flat-plate sub-results")
    ap.add_argument("--plain-stride", type=int, default=16,
                    help="ensemble: time the per-instance baseline on every k-th instance and scale by work")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: this build has no multi-GPU path (asked for --gpus {args.gpus}).",
              file=sys.stderr)
        sys.exit(2)

    if args.workload == "ensemble":
        return bench_ensemble(args)
    if args.workload == "flat_plate":
        return bench_flat_plate(args)
    if args.workload == "curvilinear":
        return bench_curvilinear(args)
    if args.workload == "cpu":
        return bench_cpu(args)
    if args.workload == "adjoint":
        out = adjoint_ratio()
        out["weno_z"] = adjoint_ratio(conv="weno_z")
        out["central_viscous"] = adjoint_ratio(viscous=True)
        print(json.dumps(out), flush=True)
        return

    import torch
    from opensbli_amd.native import Block
    from opensbli_amd.problems import taylor_green

    n = args.n
    problem, q0 = taylor_green(n, conv=args.conv, rk=args.rk)
    blk = Block(problem)
    blk.set_state(q0)
    stages = blk.num_stages()
    _log(f"bench: TGV {n}^3 {args.conv} {args.rk}, warmup {args.warmup} steps")
    blk.run(args.warmup)
    blk.sync()

    blk.profile(True)
    torch.cuda.synchronize()
    blk.sync()
    t0 = time.perf_counter()
    blk.run(args.steps)
    blk.sync()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    stats = blk.kernel_stats()
    blk.profile(False)
    work = blk.kernel_work()
    elapsed = t1 - t0
    cells = n ** 3
    value = cells * stages * args.steps / elapsed / 1e6
    ms_per_step = elapsed * 1e3 / args.steps
    stage_bytes, stage_flops = blk.work_per_stage()
    stage_evals = stages * args.steps

    # dominant kernel family by total time
    fam = max(stats, key=lambda k: stats[k][0])
    tot_ms, launches = stats[fam]
    avg_s = tot_ms / 1e3 / max(launches, 1)
    kb, kf = work[fam]
    traffic = measured_traffic(fam, n, args.conv)
    # the binding roofline of the dominant kernel: FP64 issue when its FP64
    # operations per algorithmic byte exceed the machine balance
    # 78.6e12 / 8e12 ~ 9.8 (the characteristic sweeps do ~20), HBM otherwise
    # (the central schemes' kernels do ~2-4)
    roofline_fp64 = {
        "kernel": fam,
        "bound": "fp64",
        "achieved": kf / avg_s / 1e12,
        "peak": FP64_PEAK_TFS,
        "unit": "TFLOP/s",
        "frac": kf / avg_s / 1e12 / FP64_PEAK_TFS,
        "traffic": traffic,
        "flops_per_launch": kf,
        "avg_launch_ms": avg_s * 1e3,
        "launches": launches,
    }
    roofline_hbm = {
        "kernel": fam,
        "bound": "hbm",
        "achieved": kb / avg_s / 1e9,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": kb / avg_s / 1e9 / HBM_PEAK_GBS,
        "traffic": traffic,
        "bytes_per_launch": kb,
    }
    fp64_bound = kf / max(kb, 1.0) > FP64_PEAK_TFS * 1e12 / (HBM_PEAK_GBS * 1e9)
    roofline = roofline_fp64 if fp64_bound else roofline_hbm
    stage_roofline = {
        "hbm_frac": stage_bytes * stage_evals / elapsed / 1e9 / HBM_PEAK_GBS,
        "fp64_frac": stage_flops * stage_evals / elapsed / 1e12 / FP64_PEAK_TFS,
        "bytes_per_cell_stage": stage_bytes / cells,
        "flops_per_cell_stage": stage_flops / cells,
    }
    kernels = {k: {"total_ms": v[0], "launches": v[1]} for k, v in stats.items() if v[1]}
    blk.close()

    adj = None
    if not args.no_adjoint:
        _log("bench: adjoint/forward ratio (2-D Kelvin-Helmholtz 512^2, 200 steps, checkpoints every 20)")
        adj = adjoint_ratio()
        adj["weno_z"] = adjoint_ratio(conv="weno_z")
        adj["central_viscous"] = adjoint_ratio(viscous=True)

    ens = fp = None
    if not args.no_sub:
        _log("bench: batched ensemble (512 uneven 1-D WENO5-Z shock tubes, seed 3)")
        ens = ensemble_result(args.plain_stride)
        _log("bench: flat plate 480x256x128 (tanh wall-normal mesh, factor 5)")
        fp = flat_plate_result(args.steps, args.warmup)

    cpu = None
    if not args.no_cpu_baseline:
        _log(f"bench: cpu baseline ({args.cpu_steps} step of {n}^3 on the host)")
        cpu = cpu_baseline(n, args.conv, args.cpu_steps)

    out = {
        "metric": "Mcell-updates/s per RK stage (FP64, 256^3 TGV)",
        "value": value,
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (analytic Taylor-Green initial state)",
        "config": {"workload": f"taylor_green_vortex_{n}^3_{args.conv}_roe_viscous_{args.rk}", "grid": [n, n, n],
                   "scheme": args.conv, "rk_stages": stages, "Re": 1600.0, "Minf": 0.1},
        "roofline": roofline,
        "roofline_hbm": roofline_hbm,
        "roofline_fp64": roofline_fp64,
        "stage_roofline": stage_roofline,
        "kernels": kernels,
        "ensemble": ens,
        "flat_plate": fp,
        "adjoint": adj,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def adjoint_ratio(n=512, steps=200, every=20, reps=5, conv="central_skew", viscous=False):
    """Backward/forward time of the differentiable 2-D Kelvin-Helmholtz run:
    forward = the checkpointed forward integration, backward = recomputation
    of every segment plus the hand-written adjoint steps (median of reps).
    viscous: with constant-viscosity terms (Re 1000, Pr 0.72, Minf 0.5)."""
    import numpy as np
    import torch
    from opensbli_amd.adjoint import integrate, kh_initial_state, kinetic_energy
    from opensbli_amd.native import Block
    from opensbli_amd.problems import kelvin_helmholtz
    amps = np.random.RandomState(4).normal(0.0, 1e-2, 16)
    problem, _ = kelvin_helmholtz(n, amps, conv=conv)
    if viscous:
        problem.update(viscous="const", Re=1000.0, Pr=0.72, Minf=0.5)
    b = Block(problem)
    fw, bw = [], []
    for rep in range(reps + 1):
        a = torch.tensor(amps, dtype=torch.float64, device="cuda", requires_grad=True)
        q = kh_initial_state(a, n)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        J = kinetic_energy(integrate(q, b, steps, every), problem)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        J.backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        if rep:  # the first repetition warms up
            fw.append(t1 - t0)
            bw.append(t2 - t1)
    b.close()
    f, w = float(np.median(fw)), float(np.median(bw))
    tag = "_viscous" if viscous else ""
    return {"workload": f"kelvin_helmholtz_{n}^2_{conv}{tag}_rk3_{steps}_steps_checkpoint_{every}",
            "forward_ms": f * 1e3, "backward_ms": w * 1e3, "ratio": w / f}


def flat_plate_result(steps, warmup):
    """The 480x256x128 flat plate (WENO5-Z + Sutherland viscous terms + the
    metric of the tanh-stretched wall-normal direction, RungeKuttaLS(3)):
    Mcell-updates/s per RK stage over `steps` steps and per-kernel times."""
    from opensbli_amd.native import Block
    from opensbli_amd.problems import flat_plate
    problem, q0 = flat_plate()
    blk = Block(problem)
    blk.set_state(q0)
    del q0
    stages = blk.num_stages()
    blk.run(warmup)
    blk.sync()
    blk.profile(True)
    blk.sync()
    t0 = time.perf_counter()
    blk.run(steps)
    blk.sync()
    t1 = time.perf_counter()
    stats = blk.kernel_stats()
    blk.profile(False)
    finite = bool(np.isfinite(blk.get_state()[0]).all())
    blk.close()
    n = problem["n"]
    cells = int(np.prod(n))
    el = t1 - t0
    return {
        "metric": "Mcell-updates/s per RK stage (FP64, 480x256x128 flat plate)",
        "value": cells * stages * steps / el / 1e6,
        "unit": "Mcell-updates/s",
        "steps": steps,
        "ms_per_step": el * 1e3 / steps,
        "state_finite": finite,
        "data": "synthetic (similarity boundary-layer profile + N(0,1e-4) velocity noise, seed 5)",
        "config": {"workload": f"flat_plate_{n[0]}x{n[1]}x{n[2]}_weno_z_sutherland_{problem['stretch_kind']}"
                               f"{problem['stretch_factor']:g}_metric_ls3", "grid": n, "dt": problem["dt"]},
        "kernels": {k: {"total_ms": v[0], "launches": v[1]} for k, v in stats.items() if v[1]},
    }


def bench_flat_plate(args):
    out = flat_plate_result(args.steps, args.warmup)
    out.update(n_gpus=1, warmup=args.warmup, higher_is_better=True, dtype="f64")
    print(json.dumps(out), flush=True)


def bench_curvilinear(args):
    """Curvilinear blocks (general metric, strong conservation form, WENO5-Z +
    Roe, RungeKuttaLS(3)): the Euler wave on the wavy mesh at 1024^2 and on the
    skewed wavy 3-D mesh at 128^3; Mcell-updates/s per RK stage and per-kernel
    times."""
    from opensbli_amd.native import Block
    from opensbli_amd.problems import euler_wave_curvilinear, euler_wave_curvilinear_3d
    out = {"metric": "Mcell-updates/s per RK stage (FP64, curvilinear Euler wave)", "unit": "Mcell-updates/s",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "higher_is_better": True, "dtype": "f64",
           "data": "synthetic (analytic density wave on analytic meshes)", "cases": {}}
    for name, (problem, q0) in (("wavy_2d_1024^2", euler_wave_curvilinear(1024)),
                                ("skewed_wavy_3d_128^3", euler_wave_curvilinear_3d(128))):
        _log(f"bench: curvilinear {name}")
        blk = Block(problem)
        blk.set_state(q0)
        stages = blk.num_stages()
        blk.run(args.warmup)
        blk.sync()
        blk.profile(True)
        blk.sync()
        t0 = time.perf_counter()
        blk.run(args.steps)
        blk.sync()
        t1 = time.perf_counter()
        stats = blk.kernel_stats()
        blk.profile(False)
        blk.close()
        cells = int(np.prod(problem["n"]))
        el = t1 - t0
        out["cases"][name] = {
            "value": cells * stages * args.steps / el / 1e6,
            "ms_per_step": el * 1e3 / args.steps,
            "kernels": {k: {"total_ms": v[0], "launches": v[1]} for k, v in stats.items() if v[1]},
        }
    print(json.dumps(out), flush=True)


def ensemble_result(plain_stride=16):
    """512 uneven 1-D shock tubes (WENO5-Z, RK3, CFL 0.5, t = 0.2): one
    persistent launch vs a graph-staged launch sequence vs per-instance runs
    (timed on every plain_stride-th instance and scaled by its share of the
    cell updates)."""
    import torch
    from opensbli_amd.ensemble import Ensemble, autotune, ensemble_parameters, plain_runs

    par = ensemble_parameters(512, seed=3)
    t0 = time.perf_counter()
    max_n, threads = autotune(par)
    tune_s = time.perf_counter() - t0
    e = Ensemble(par)
    e.set_threads(max_n, threads)
    e.save()
    e.run("persistent")  # warm-up
    e.restore()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ms_p = e.run("persistent")
    t_p = time.perf_counter() - t0
    e.restore()
    ms_s = e.run("staged")
    cu = e.cell_updates()
    e.close()
    idx = list(range(0, 512, plain_stride))
    t_plain_sample = plain_runs(par, indices=idx)
    work = par["n"].astype(float) * par["nsteps"]
    frac = work[idx].sum() / work.sum()
    t_plain = t_plain_sample / frac
    return {
        "metric": "batched ensemble speed-up over per-instance launches (512 1-D WENO5-Z shock tubes)",
        "value": t_plain / (ms_p / 1e3),
        "unit": "x",
        "data": "synthetic (seeded shock tubes, numpy RandomState(3))",
        "config": {"workload": "ensemble_512_shock_tubes_weno_z_rk3", "instances": 512,
                   "n_range": [int(par["n"].min()), int(par["n"].max())], "t_end": 0.2, "cfl": 0.5},
        "persistent_ms": ms_p,
        "persistent_wall_s": t_p,
        "staged_graph_ms": ms_s,
        "plain_s_estimate": t_plain,
        "plain_sample": f"every {plain_stride}th instance ({len(idx)} runs, {frac:.3f} of the work) "
                        f"took {t_plain_sample:.2f} s",
        "cell_updates": cu,
        "mcell_updates_per_s_persistent": cu / ms_p / 1e3,
        "mcell_updates_per_s_staged": cu / ms_s / 1e3,
        "autotune": {"max_n": max_n, "threads": threads, "seconds": tune_s},
    }


def bench_ensemble(args):
    out = ensemble_result(args.plain_stride)
    out.update(n_gpus=1, higher_is_better=True, dtype="f64")
    print(json.dumps(out), flush=True)


def _oracle_instance(i, par):
    """Run ensemble instance i on the oracle to its step count (process-pool worker)."""
    from opensbli_amd.ensemble import instance_problem
    from oracle.ref import RefSolver
    problem, q0 = instance_problem(par, i)
    r = RefSolver(problem, variant="fma")
    r.set_state(q0)
    r.run(int(par["nsteps"][i]))
    return i


def bench_cpu(args):
    """The oracle on this host for the other configurations, each as one
    warm-up run then the median of three timed runs: the 64^3 central TGV
    (RungeKuttaLS(3), 20 steps), the 512^2 Kelvin-Helmholtz layer (central
    skew, RK3, 20 steps), the 480x256x128 flat plate (1 step) and the
    512-tube ensemble (every 8th instance to t = 0.2, one single-threaded
    oracle process per core)."""
    from opensbli_amd.ensemble import ensemble_parameters
    from opensbli_amd.problems import flat_plate, kelvin_helmholtz, taylor_green
    from oracle.ref import RefSolver
    threads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
    out = {"metric": "oracle Mcell-updates/s per RK stage on the host", "unit": "Mcell-updates/s",
           "cores": threads, "kind": "port", "host": host_cpu(),
           "build": "-O3 -march=x86-64-v3 -fopenmp -ffp-contract=fast", "cases": {}}

    def timed_case(name, problem, q0, steps):
        _log(f"bench: cpu {name}")
        s = RefSolver(problem, variant="fma")
        s.set_state(q0)
        dt, ts = _median_timed(s, steps)
        cells = int(np.prod(problem["n"][: problem["ndim"]]))
        stages = {"rk3": 3, "ls3": 3, "ls3_ssp": 3, "ls4": 4}[problem.get("rk", "rk3")]
        out["cases"][name] = {"value": cells * stages * steps / dt / 1e6, "steps": steps,
                              "runs_s": [round(t, 3) for t in ts]}

    p, q = taylor_green(64, conv="central_skew", rk="ls3")
    timed_case("taylor_green_64^3_central_skew_ls3", p, q, 20)
    amps = np.random.RandomState(4).normal(0.0, 1e-2, 16)
    p, q = kelvin_helmholtz(512, amps, conv="central_skew")
    timed_case("kelvin_helmholtz_512^2_central_skew_rk3", p, q, 20)
    p, q = flat_plate()
    timed_case("flat_plate_480x256x128_weno_z_sutherland_tanh5_ls3", p, q, 1)
    del q
    # the tubes are independent: one single-threaded oracle process per core,
    # instances handed out in order of decreasing work
    import multiprocessing as mp
    from concurrent.futures import ProcessPoolExecutor
    par = ensemble_parameters(512, seed=3)
    idx = list(range(0, 512, 8))
    work = par["n"].astype(float) * par["nsteps"]
    idx.sort(key=lambda i: -work[i])
    _log(f"bench: cpu ensemble sample ({len(idx)} instances on {threads} single-threaded processes)")
    omp = os.environ.get("OMP_NUM_THREADS")
    os.environ["OMP_NUM_THREADS"] = "1"
    try:
        with ProcessPoolExecutor(max_workers=threads, mp_context=mp.get_context("spawn")) as ex:
            list(ex.map(_oracle_instance, [0], [par]))  # start the workers
            t0 = time.perf_counter()
            list(ex.map(_oracle_instance, idx, [par] * len(idx)))
            el = time.perf_counter() - t0
    finally:
        if omp is None:
            del os.environ["OMP_NUM_THREADS"]
        else:
            os.environ["OMP_NUM_THREADS"] = omp
    out["cases"]["ensemble_512_shock_tubes_weno_z_rk3"] = {
        "value": 3 * work[idx].sum() / el / 1e6, "sample": f"every 8th instance ({len(idx)} of 512, "
        f"{work[idx].sum() / work.sum():.3f} of the cell updates) to t = 0.2 on {threads} single-threaded oracle "
        f"processes in {el:.1f} s", "estimated_all_512_s": el * work.sum() / work[idx].sum()}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
