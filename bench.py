#!/usr/bin/env python3
"""Benchmark: cell-steps/s of the full solver step (MUSCL transport + MGPCG projection to
rtol 1e-8 + surface loads) on one MI355X.

Default workload: NACA 2412 wing, alpha 4 deg, 2-level static AMR of 8^3 bricks (~4.1 M
active cells), FP32 fields with the FP64 outer conjugate-gradient solve.  A "step" is one
call of the solver step on the resident state; the loads integral runs once per step too.
Prints one JSON line (rank 0).  --gpus above 1 is not supported on this single-GPU build.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "cell-steps/s (MUSCL+MGPCG to rtol 1e-8), learned vs plain PC, 1 MI355X"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(topo, options, budget_s):
    """Serial FP64 oracle (one host core) on a bounded sample of the same workload."""
    from oracle.oracle import Oracle, build
    from paracfd_amd import _lib as L
    build()
    o = L.Options()
    for name, _ in L.Options._fields_:
        setattr(o, name, getattr(options, name))
    o.field_precision = L.FP64
    o.vcycle_precision = L.FP64
    orc = Oracle(topo, o)
    t0 = time.perf_counter()
    orc.initialize()
    t_init = time.perf_counter() - t0
    log(f"cpu oracle initialize {t_init:.1f}s")
    steps = 0
    elapsed = 0.0
    while steps == 0 or elapsed < budget_s:
        elapsed += orc.time_steps(1)
        steps += 1
        log(f"cpu oracle step {steps}: {elapsed:.1f}s")
        if elapsed + elapsed / steps > 2 * budget_s:
            break
    orc.close()
    n = topo.n_active
    return {"value": n * steps / elapsed, "unit": "cell-steps/s", "cores": 1, "kind": "port",
            "sample": f"{steps} step(s) after initialization of the same {n}-cell case, serial FP64 oracle "
                      f"({elapsed:.1f} s timed)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--case", default="naca2412_2lvl")
    ap.add_argument("--fields", default="fp32", choices=["fp32", "fp64"])
    ap.add_argument("--vcycle", default="fp64", choices=["fp64", "fp32", "bf16"])
    ap.add_argument("--check-interval", type=int, default=2)
    ap.add_argument("--cpu-budget", type=float, default=15.0, help="seconds of oracle time for the CPU baseline")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs on one GPU only; --gpus must be 1 (multi-GPU sharding is designed, "
              "not built; see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import torch
    if not torch.cuda.is_available():
        print("bench.py: no GPU visible", file=sys.stderr)
        sys.exit(1)

    from paracfd_amd import ExternalAeroCore, _lib as L, cases, default_options

    g, d, spec = cases.build_case(args.case)
    t0 = time.perf_counter()
    topo = L.Topology(g, d)
    log(f"topology {args.case}: {topo.n_active} active cells, {topo.info.bricks} bricks, "
        f"{topo.mg_levels} MG levels, built in {time.perf_counter() - t0:.1f}s")
    prec = {"fp64": L.FP64, "fp32": L.FP32, "bf16": L.BF16}
    opt = default_options(rho=spec["rho"], nu=spec["nu"], speed=spec["speed"], cfl=spec["cfl"], rtol=1e-8,
                          max_iterations=500, field_precision=prec[args.fields], vcycle_precision=prec[args.vcycle],
                          preconditioner=L.PC_MG, pre_smooth=2, post_smooth=2, check_interval=args.check_interval)
    core = ExternalAeroCore(topo, opt)
    core.set_reference(topo.info.reference_area)
    st = core.initialize()
    log(f"initialize: {st['iterations']} PCG iterations, relres {st['relative_residual']:.2e}, "
        f"converged {st['converged']}")
    for i in range(args.warmup):
        st = core.step()
        core.pressure_loads()
        log(f"warmup {i}: dt {st['dt']:.4g} its {st['iterations']} step {st['step_ms']:.2f} ms")
    torch.cuda.synchronize()
    its = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        st = core.step()
        loads = core.pressure_loads()
        its.append(st["iterations"])
        if not st["converged"]:
            log(f"step {i} did not converge: relres {st['relative_residual']:.3e}")
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    seconds = t1 - t0
    n = topo.n_active
    value = n * args.steps / seconds
    log(f"timed {args.steps} steps in {seconds:.3f}s: {value:.4g} cell-steps/s, mean its {sum(its) / len(its):.1f}, "
        f"CL {loads['cl']:.4f} CD {loads['cd']:.4f}")

    # roofline of the dominant kernel, timed live with HIP events on the solver's stream
    mean_its = sum(its) / len(its)
    # launches per step of each timed kernel family: the CG direction+apply pass once per
    # iteration; four composite-level sweep-type passes per V-cycle (2 pre-sweeps fused,
    # residual+restriction, prolongation+sweep, last sweep), one V-cycle per iteration plus
    # the initial one; one transport pass
    per_step = {
        0: mean_its,
        1: (mean_its + 1) * 4,
        2: 1.0,
    }
    timings = {}
    for k in per_step:
        ms, by = core.time_kernel(k, reps=20)
        timings[k] = (ms, by, ms * per_step[k])
        log(f"kernel {k}: {ms * 1e3:.1f} us/launch, {by / (ms * 1e-3) / 1e9:.0f} GB/s algorithmic, "
            f"{ms * per_step[k]:.2f} ms/step")
    names = {0: "k_cg_apply_dot", 1: "k_smooth0", 2: "k_advect"}
    dom = max(timings, key=lambda k: timings[k][2])
    ms, by, _ = timings[dom]
    achieved = by / (ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "kernel": names[dom], "achieved": achieved, "peak": 8000.0, "unit": "GB/s",
                "frac": achieved / 8000.0, "traffic": None}

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(topo, opt, args.cpu_budget)
        log(f"cpu baseline {cpu['value']:.4g} cell-steps/s")

    out = {
        "metric": METRIC, "value": value, "unit": "cell-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * seconds / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 fields + f64 PCG" if args.fields == "fp32" else "f64",
        "data": "synthetic (analytic geometry, freestream start)",
        "config": {"workload": args.case, "active_cells": n, "amr_levels": topo.info.levels,
                   "preconditioner": "geometric V-cycle MGPCG", "vcycle": args.vcycle, "rtol": 1e-8,
                   "mean_pcg_iterations": mean_its, "cl": loads["cl"], "cd": loads["cd"]},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
