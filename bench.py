#!/usr/bin/env python
"""Throughput of the OpenSBLI RHS + RK3 time advance on one MI355X.

Workload: 3-D Taylor-Green vortex, 256^3 periodic, FP64, Re = 1600, M = 0.1,
WENO5-Z characteristic LLF flux (Roe average) with 4th-order central viscous
terms and the 3-register RK3 (the TGV script's RungeKutta(3)).  A step is one
RK3 iteration (three stages).  The state is resident in HBM before timing.

Prints one JSON line:
  value = Mcell-updates/s per RK stage = cells * stages * steps / t / 1e6
  roofline = the dominant kernel's algorithmic bytes per launch / its mean
             launch time (HIP events on the library's stream) vs 8 TB/s
  cpu_baseline = the oracle (C restatement of the generated OPS kernel
             sequence) on the host's cores, one step of the same workload.
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
# the commands and raw figures are in profiles/hbm_traffic_weno_z_tgv256.md
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic_weno_z_tgv256.json")
FP64_PEAK_TFS = 78.6


def _log(msg):
    print(msg, file=sys.stderr, flush=True)


def measured_traffic(family, n, conv):
    """HBM bytes per launch of a kernel family from the committed measurement,
    when it covers this workload; None otherwise."""
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    if t.get("grid") != n or t.get("scheme") != conv:
        return None
    fam = t.get("families", {}).get(family)
    return None if fam is None else fam.get("bytes_per_launch")


def cpu_baseline(n, conv, steps=1):
    """Time the oracle on this host, the same problem, `steps` RK3 steps."""
    from opensbli_amd.problems import taylor_green
    from oracle.ref import RefSolver
    threads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
    problem, q0 = taylor_green(n, conv=conv)
    s = RefSolver(problem)
    s.set_state(q0)
    t0 = time.perf_counter()
    s.run(steps)
    dt = time.perf_counter() - t0
    return {
        "value": n ** 3 * 3 * steps / dt / 1e6,
        "unit": "Mcell-updates/s per RK stage",
        "cores": threads,
        "kind": "port",
        "sample": f"{steps} RK3 step(s) of the same {n}^3 TGV {conv} problem "
                  f"({dt:.1f} s, OMP_NUM_THREADS={threads})",
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
    ap.add_argument("--workload", default="tgv", choices=["tgv", "ensemble", "flat_plate", "adjoint",
                                                           "curvilinear"])
    ap.add_argument("--no-adjoint", action="store_true", help="tgv: skip the adjoint/forward ratio measurement")
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
    if args.workload == "adjoint":
        out = adjoint_ratio()
        out["weno_z"] = adjoint_ratio(conv="weno_z", reps=1)
        out["central_viscous"] = adjoint_ratio(viscous=True, reps=1)
        print(json.dumps(out), flush=True)
        return

    import torch
    from opensbli_amd.native import Block
    from opensbli_amd.problems import taylor_green

    n = args.n
    problem, q0 = taylor_green(n, conv=args.conv)
    blk = Block(problem)
    blk.set_state(q0)
    stages = blk.num_stages()
    _log(f"bench: TGV {n}^3 {args.conv} RK3, warmup {args.warmup} steps")
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
    roofline = {
        "kernel": fam,
        "bound": "hbm",
        "achieved": kb / avg_s / 1e9,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": kb / avg_s / 1e9 / HBM_PEAK_GBS,
        "traffic": measured_traffic(fam, n, args.conv),
        "avg_launch_ms": avg_s * 1e3,
        "launches": launches,
    }
    roofline_fp64 = {
        "kernel": fam,
        "achieved": kf / avg_s / 1e12,
        "peak": FP64_PEAK_TFS,
        "unit": "TFLOP/s",
        "frac": kf / avg_s / 1e12 / FP64_PEAK_TFS,
    }
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
        adj["weno_z"] = adjoint_ratio(conv="weno_z", reps=1)
        adj["central_viscous"] = adjoint_ratio(viscous=True, reps=1)

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
        "config": {"workload": f"taylor_green_vortex_{n}^3_{args.conv}_roe_viscous_rk3", "grid": [n, n, n],
                   "scheme": args.conv, "rk_stages": stages, "Re": 1600.0, "Minf": 0.1},
        "roofline": roofline,
        "roofline_fp64": roofline_fp64,
        "stage_roofline": stage_roofline,
        "kernels": kernels,
        "adjoint": adj,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def adjoint_ratio(n=512, steps=200, every=20, reps=3, conv="central_skew", viscous=False):
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


def bench_flat_plate(args):
    """The 480x256x128 flat plate (WENO5-Z + Sutherland viscous + metric,
    RungeKuttaLS(3)): Mcell-updates/s per RK stage and per-kernel times."""
    from opensbli_amd.native import Block
    from opensbli_amd.problems import flat_plate
    problem, q0 = flat_plate()
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
    cells = 480 * 256 * 128
    el = t1 - t0
    out = {
        "metric": "Mcell-updates/s per RK stage (FP64, 480x256x128 flat plate)",
        "value": cells * stages * args.steps / el / 1e6,
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el * 1e3 / args.steps,
        "higher_is_better": True,
        "dtype": "f64",
        "data": "synthetic (similarity boundary-layer profile + N(0,1e-4) velocity noise, seed 5)",
        "config": {"workload": "flat_plate_480x256x128_weno_z_sutherland_metric_ls3", "grid": [480, 256, 128]},
        "kernels": {k: {"total_ms": v[0], "launches": v[1]} for k, v in stats.items() if v[1]},
    }
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


def bench_ensemble(args):
    """512 uneven 1-D shock tubes (WENO5-Z, RK3, CFL 0.5, t = 0.2): one
    persistent launch vs a graph-staged launch sequence vs per-instance runs."""
    import numpy as np
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
    idx = list(range(0, 512, args.plain_stride))
    t_plain_sample = plain_runs(par, indices=idx)
    work = par["n"].astype(float) * par["nsteps"]
    frac = work[idx].sum() / work.sum()
    t_plain = t_plain_sample / frac
    out = {
        "metric": "batched ensemble speed-up over per-instance launches (512 1-D WENO5-Z shock tubes)",
        "value": t_plain / (ms_p / 1e3),
        "unit": "x",
        "n_gpus": 1,
        "higher_is_better": True,
        "dtype": "f64",
        "data": "synthetic (seeded shock tubes, numpy RandomState(3))",
        "config": {"workload": "ensemble_512_shock_tubes_weno_z_rk3", "instances": 512,
                   "n_range": [int(par["n"].min()), int(par["n"].max())], "t_end": 0.2, "cfl": 0.5},
        "persistent_ms": ms_p,
        "persistent_wall_s": t_p,
        "staged_graph_ms": ms_s,
        "plain_s_estimate": t_plain,
        "plain_sample": f"every {args.plain_stride}th instance ({len(idx)} runs, {frac:.3f} of the work) "
                        f"took {t_plain_sample:.2f} s",
        "cell_updates": cu,
        "mcell_updates_per_s_persistent": cu / ms_p / 1e3,
        "mcell_updates_per_s_staged": cu / ms_s / 1e3,
        "autotune": {"max_n": max_n, "threads": threads, "seconds": tune_s},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
