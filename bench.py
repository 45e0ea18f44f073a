#!/usr/bin/env python3
"""Benchmark: cell-steps/s of the full solver step (MUSCL transport + MGPCG projection to
rtol 1e-8 + surface loads) on one MI355X, learned preconditioner vs plain MGPCG.

Default workload: NACA 2412 wing, alpha 4 deg, 3-level static AMR of 8^3 bricks (19,840,144
active cells, refinement bands 1.0 and 0.25 chord), FP32 fields, FP64 outer conjugate
gradients.  Three solvers run the same step sequence from the same freestream start:
  * learned: FP32 V-cycle + the learned network correction (BF16 MFMA) in flexible PCG, after
    untimed data collection and training on the solver's own MGPCG residuals  -> `value`
  * plain:   FP32 V-cycle MGPCG (the same arithmetic without the network)
  * plain_fp64_vcycle: FP64 V-cycle MGPCG (plain FP64-outer MGPCG, the reference's method)
and an FP64-everywhere solver (fields, V-cycle, CG; it matches the CPU oracle to 1e-10 in the
parity tests) gives the errors of the learned run's velocity, pressure and CL/CD.  Two more
runs solve the full operator with the non-orthogonal correction (FP32 V-cycle; the reference's
default semantics) and report their throughput beside: nonorthogonal_fgmres (plain flexible
GMRES) and nonorthogonal_gcro (GCRO with three recycled corrections, as the reference).
A "step" is one solver step on the resident state plus the loads integral.  Prints one JSON
line.  `roofline` is the hot kernel with the largest time per step, timed in-step: HIP events
around each of its launches on the solver's stream during steps that follow the timed ones
(pcfd_probe_start/stop; the iterations are then launched one by one instead of replayed from
graphs).  The timed steps and the probe steps sit in roctx ranges ("timed:<solver>",
"probe:<solver>") for rocprofv3 --marker-trace.  --gpus above 1 is not supported on this single-GPU build.
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
PEAK_HBM = 8000.0      # GB/s
PEAK_BF16 = 2500.0     # TFLOP/s dense


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


class _Roctx:
    """roctx ranges for rocprofv3 --marker-trace (no-ops when the library is absent)."""

    def __init__(self):
        import ctypes
        self.lib = None
        for name in ("librocprofiler-sdk-roctx.so.1", "librocprofiler-sdk-roctx.so"):
            try:
                self.lib = ctypes.CDLL(name)
                break
            except OSError:
                continue
        if self.lib is not None:
            self.lib.roctxRangePushA.argtypes = [ctypes.c_char_p]

    def push(self, name):
        if self.lib is not None:
            self.lib.roctxRangePushA(name.encode())

    def pop(self):
        if self.lib is not None:
            self.lib.roctxRangePop()


ROCTX = _Roctx()


def host_cpu():
    """Model name and thread counts of the host the CPU baseline runs on."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = None
    return {"cpu_model": model, "host_logical_cpus": os.cpu_count(), "usable_cpus": usable}


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
    o.preconditioner = L.PC_MG
    orc = Oracle(topo, o)
    t0 = time.perf_counter()
    orc.initialize()
    log(f"cpu oracle initialize {time.perf_counter() - t0:.1f}s")
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
    out = {"value": n * steps / elapsed, "unit": "cell-steps/s", "cores": 1, "kind": "port",
           "sample": f"{steps} step(s) after initialization of the same {n}-cell case, serial FP64 oracle "
                     f"({elapsed:.1f} s timed)"}
    out.update(host_cpu())
    return out


def device_ceilings():
    """Practical ceilings measured on this GPU: a 1 GiB device-to-device copy (read + write bytes
    per second) and a 8192^3 BF16 GEMM (hipBLASLt through torch.matmul), HIP events."""
    import torch
    n = 1 << 28
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(10):
        b.copy_(a)
    e1.record()
    e1.synchronize()
    copy_gbs = 10 * 2 * 4 * n / (e0.elapsed_time(e1) * 1e-3) / 1e9
    del a, b
    m = 8192
    x = torch.randn(m, m, dtype=torch.bfloat16, device="cuda")
    y = torch.randn(m, m, dtype=torch.bfloat16, device="cuda")
    torch.matmul(x, y)
    e0.record()
    for _ in range(10):
        torch.matmul(x, y)
    e1.record()
    e1.synchronize()
    gemm = 10 * 2.0 * m ** 3 / (e0.elapsed_time(e1) * 1e-3) / 1e12
    del x, y
    torch.cuda.empty_cache()
    return {"copy_GBps": copy_gbs, "bf16_gemm_TFLOPs": gemm,
            "note": "1 GiB device copy (read + write bytes) and an 8192^3 BF16 GEMM via torch.matmul, HIP events"}


def measured_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` on `workload` from the committed PMC summary, or None."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            entries = json.load(f).get("entries", [])
    except (OSError, ValueError):
        return None
    for rec in entries:
        if rec.get("kernel") == kernel and rec.get("workload") == workload:
            return rec["hbm_bytes_per_launch"]
    return None


def rel_l2(a, b):
    import numpy as np
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def run(topo, spec, args, name, field, vcycle, pc, nonorthogonal=0, gmres_recycle=3, pretrained=None):
    """Initialise, pre-steps (with training for the learned solver), warmup, K timed steps."""
    import torch
    from paracfd_amd import ExternalAeroCore, _lib as L, default_options
    opt = default_options(rho=spec["rho"], nu=spec["nu"], speed=spec["speed"], cfl=spec["cfl"], rtol=1e-8,
                          max_iterations=500, field_precision=field, vcycle_precision=vcycle, preconditioner=pc,
                          pre_smooth=2, post_smooth=2, check_interval=args.check_interval,
                          learned_bricks=args.learned_bricks, learned_capacity=args.collects * args.pairs,
                          smooth_wall=1 if args.smooth_wall else 0, nonorthogonal=nonorthogonal,
                          gmres_recycle=gmres_recycle)
    core = ExternalAeroCore(topo, opt)
    core.set_reference(topo.info.reference_area)
    st = core.initialize()
    log(f"{name}: initialize {st['iterations']} its, converged {st['converged']}")
    if pc == L.PC_LEARNED and pretrained is not None:
        core.set_learned_weights(*pretrained)
    train = None
    for i in range(2 + args.collects):
        if pc == L.PC_LEARNED and i >= 2:
            core.learned_collect(max_pairs=args.pairs)
        core.step()
        core.pressure_loads()
    if pc == L.PC_LEARNED:
        train = core.learned_train(steps=args.train_steps, batch=4, lr=args.lr)
        log(f"{name}: trained {train['train_steps']} steps on {train['pairs']} residuals "
            f"({train['bricks']} bricks), loss {train['loss_first']:.4f} -> {train['loss_last']:.4f}, "
            f"{train['train_ms']:.0f} ms")
    # every solver runs the same steps; for the learned one they are the validation that
    # alternates learned and plain projections (6 steps)
    trial = []
    for i in range(6):
        torch.cuda.synchronize()
        t = time.perf_counter()
        st = core.step()
        core.pressure_loads()
        torch.cuda.synchronize()
        trial.append(f"{'learned' if st['learned_used'] else 'plain'} {st['iterations']} its "
                     f"{1e3 * (time.perf_counter() - t):.1f} ms")
    if pc == L.PC_LEARNED:
        log(f"{name}: validation steps: " + "; ".join(trial))
        log(f"{name}: network {'rejected' if st['learned_rejected'] else 'kept'} by the validation")
    tune = core.autotune(repetitions=2) if args.autotune else core.tuning()
    log(f"{name}: launch shapes {tune}")
    for i in range(args.warmup):
        st = core.step()
        core.pressure_loads()
    torch.cuda.synchronize()
    its, fallbacks, used, conv = [], 0, 0, True
    ROCTX.push(f"timed:{name}")
    t0 = time.perf_counter()
    for i in range(args.steps):
        st = core.step()
        loads = core.pressure_loads()
        its.append(st["iterations"])
        fallbacks += st["fallback_used"]
        used += st["learned_used"]
        conv = conv and bool(st["converged"])
    torch.cuda.synchronize()
    seconds = time.perf_counter() - t0
    ROCTX.pop()
    n = topo.n_active
    res = {"core": core, "value": n * args.steps / seconds, "ms_per_step": 1e3 * seconds / args.steps,
           "mean_its": sum(its) / len(its), "fallback_steps": fallbacks, "converged": conv, "loads": loads,
           "train": train, "learned_steps": used, "learned_rejected": bool(st["learned_rejected"]), "tuning": tune,
           "substeps": st["diffusion_substeps"]}
    log(f"{name}: {res['value']:.4g} cell-steps/s, {res['ms_per_step']:.2f} ms/step, mean its {res['mean_its']:.1f}, "
        f"fallbacks {fallbacks}, network used in {used} steps (rejected: {res['learned_rejected']}), "
        f"CL {loads['cl']:.6f} CD {loads['cd']:.6f}")
    return res


def pretrain_network(args, spec, vc):
    """Weights of the network trained across args.pretrain angles of attack in [0, 12] deg on
    the case's geometry at base resolution args.pretrain_cells_x (None without pretraining)."""
    if args.pretrain <= 0 or args.case not in ("naca2412_3lvl", "naca2412_2lvl", "paraglider_3lvl"):
        return None, None
    from paracfd_amd import _lib as L, cases, default_options, learned as LN
    opt = default_options(rho=spec["rho"], nu=spec["nu"], speed=spec["speed"], cfl=spec["cfl"], rtol=1e-8,
                          max_iterations=500, field_precision=L.FP32, vcycle_precision=vc,
                          preconditioner=L.PC_LEARNED, pre_smooth=2, post_smooth=2,
                          check_interval=args.check_interval, learned_bricks=args.learned_bricks,
                          learned_capacity=args.collects * args.pairs)

    def make(alpha, device):
        g, d, _ = cases.build_case(args.case, cells_x=args.pretrain_cells_x, alpha_deg=alpha)
        return L.Topology(g, d, device=device)

    angles = LN.default_angles(args.pretrain)
    W, B, stats = LN.train_across_angles(make, angles, opt, collects=args.collects, pairs=args.pairs,
                                         train_steps=args.train_steps, lr=args.lr, log=log)
    log(f"pretraining: {len(angles)} angles in [0, 12] deg, {stats['residuals']} residuals, "
        f"{stats['brick_samples']} brick samples, {stats['seconds']:.1f}s")
    summary = {"angles_deg": [round(a, 3) for a in angles], "cells_x": args.pretrain_cells_x,
               "residuals": stats["residuals"], "brick_samples": stats["brick_samples"],
               "seconds": round(stats["seconds"], 2),
               "loss_last": [round(r["loss_last"], 4) for r in stats["instances"]]}
    return (W, B), summary


def run_sweep(args):
    """Batched angle sweep: every instance of the sweep in one solver on the merged topology
    (one step advances all of them), against the same instances stepped one after another in
    their own solvers.  Mixed precision: FP32 fields and V-cycle, FP64 outer PCG and loads."""
    import numpy as np
    import torch
    from paracfd_amd import ExternalAeroCore, _lib as L, cases, default_options

    t0 = time.perf_counter()
    parts, spec = cases.build_sweep(args.case, device=None if args.host_topology else 0)
    merged = L.Topology.merge(parts)
    n_total = sum(p.n_active for p in parts)
    log(f"sweep {args.case}: {len(parts)} instances, {n_total} active cells "
        f"({min(p.n_active for p in parts)}..{max(p.n_active for p in parts)} per instance), "
        f"{merged.info.bricks} bricks, built and merged in {time.perf_counter() - t0:.1f}s")
    opt = default_options(rho=spec["rho"], nu=spec["nu"], speed=spec["speed"], cfl=spec["cfl"], rtol=1e-8,
                          max_iterations=500, field_precision=L.FP32, vcycle_precision=L.FP32,
                          preconditioner=L.PC_MG, pre_smooth=2, post_smooth=2, check_interval=args.check_interval,
                          smooth_wall=1 if args.smooth_wall else 0)

    def advance(core, topo, loads_fn):
        core.initialize()
        core.step()
        tune = core.autotune(repetitions=2) if args.autotune else core.tuning()
        for _ in range(1 + args.warmup):
            core.step()
            loads_fn()
        torch.cuda.synchronize()
        its = []
        t = time.perf_counter()
        for _ in range(args.steps):
            st = core.step()
            loads = loads_fn()
            its.append(st["iterations"])
        torch.cuda.synchronize()
        return time.perf_counter() - t, its, loads, tune

    batch = ExternalAeroCore(merged, opt)
    for i, p in enumerate(parts):
        batch.set_instance_reference(i, p.info.reference_area)
    log(f"batch solver: {batch.device_bytes() / 2**30:.1f} GiB of device memory")
    tb, its_b, loads_b, tune_b = advance(batch, merged, batch.batch_loads)
    log(f"batch launch shapes {tune_b}")
    per_inst = [batch.instance_stats(i) for i in range(len(parts))]
    value = n_total * args.steps / tb
    log(f"batch: {value:.4g} cell-steps/s, {1e3 * tb / args.steps:.1f} ms/step, mean its {np.mean(its_b):.1f}")
    ms_apply, by_apply = batch.time_kernel(6, reps=20)   # the dominant kernel of a PCG iteration
    batch.close()

    seq_s, seq_loads, seq_its = 0.0, [], []
    for i, p in enumerate(parts):
        core = ExternalAeroCore(p, opt)
        core.set_reference(p.info.reference_area)
        t, its, loads, _ = advance(core, p, core.pressure_loads)
        core.close()
        seq_s += t
        seq_loads.append(loads)
        seq_its.append(float(np.mean(its)))
        log(f"instance {i} alone: {p.n_active * args.steps / t:.4g} cell-steps/s, mean its {seq_its[-1]:.1f}, "
            f"CL {loads['cl']:.6f} CD {loads['cd']:.6f}")
    seq_value = n_total * args.steps / seq_s
    dl = max(max(abs(a["cl"] - b["cl"]) / max(abs(b["cl"]), 1e-30), abs(a["cd"] - b["cd"]) / max(abs(b["cd"]), 1e-30))
             for a, b in zip(loads_b, seq_loads))
    achieved = by_apply / (ms_apply * 1e-3) / 1e9
    out = {
        "metric": "aggregate cell-steps/s (MUSCL+MGPCG to rtol 1e-8), batched angle sweep, 1 MI355X",
        "value": value, "unit": "cell-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * tb / args.steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 fields + f32 V-cycle + f64 PCG", "data": "synthetic (analytic geometry, freestream start)",
        "config": {"workload": args.case, "instances": len(parts), "active_cells": n_total,
                   "alphas": list(spec["alphas"]), "rtol": 1e-8, "mean_pcg_iterations_batch": float(np.mean(its_b)),
                   "sequential_value": seq_value, "speedup_vs_sequential": value / seq_value,
                   "launch_shapes": tune_b,
                   "max_rel_load_difference_batch_vs_sequential": dl,
                   "instances_detail": [{"alpha": a, "cells": p.n_active, "cl": l["cl"], "cd": l["cd"],
                                         "iterations": s["iterations"], "dt": s["dt"], "mean_its_alone": si}
                                        for a, p, l, s, si in zip(spec["alphas"], parts, loads_b, per_inst, seq_its)]},
        "roofline": {"bound": "hbm", "kernel": "k_vc_update_pre2<float>", "achieved": achieved, "peak": PEAK_HBM,
                     "unit": "GB/s", "frac": achieved / PEAK_HBM, "traffic": None},
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def kernel_probe(args):
    """A series of launches of one hot kernel (no other launch of it in the process) and plain
    device copies of known size, for rocprofv3 --pmc runs: the counters of the copies give the
    scale factor that turns FETCH_SIZE / WRITE_SIZE of the kernel into HBM bytes."""
    import torch
    from paracfd_amd import ExternalAeroCore, _lib as L, cases, default_options
    g, d, spec = cases.build_case(args.case)
    topo = L.Topology(g, d)
    core = ExternalAeroCore(topo, default_options(rho=spec["rho"], nu=spec["nu"], speed=spec["speed"], cfl=spec["cfl"],
                                                  field_precision=L.FP32, vcycle_precision=L.FP32))
    core.set_tuning(xcd_chunk_log2=args.probe_chunk)
    ms, by = core.time_kernel(args.kernel_probe, reps=args.probe_reps)
    nbytes = 256 * 2**20
    a = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"kernel_id": args.kernel_probe, "workload": args.case, "active_cells": topo.n_active,
                      "us_per_launch": ms * 1e3, "algorithmic_bytes": by, "launches": args.probe_reps + 1,
                      "copy_bytes": nbytes, "copies": 3, "xcd_chunk_log2": args.probe_chunk}), flush=True)


def topology_build(args):
    """Host build and device build of the case's topology: stage times and bit-identity of every
    exported array (topology and multigrid levels)."""
    from paracfd_amd import _lib as L, cases
    g, d, spec = cases.build_case(args.case)
    res = {}
    topos = {}
    for where, dev in (("device", 0), ("host", None)):
        log(f"{where} build of {args.case}")
        t0 = time.perf_counter()
        topos[where] = L.Topology(g, d, device=dev)
        wall = time.perf_counter() - t0
        stages, _ = topos[where].build_times()
        res[where] = {"wall_s": wall, "stages_ms": {k: round(v, 1) for k, v in stages.items()}}
    h, dv = topos["host"], topos["device"]
    identical, values = True, 0
    for key in L.ARR:
        if key.startswith(("nf_", "ng_")):
            continue
        a, b = h.array(key), dv.array(key)
        identical &= a.dtype == b.dtype and a.tobytes() == b.tobytes()
        values += a.size
    for level in range(h.mg_levels):
        for key in L.MG_ARR:
            try:
                a = h.mg_array(level, key)
            except KeyError:
                continue
            identical &= a.tobytes() == dv.mg_array(level, key).tobytes()
            values += a.size
    print(json.dumps({"workload": args.case, "active_cells": h.n_active, "special_cells": int(h.info.special_cells),
                      "compact_edges": int(h.info.compact_edges), "bit_identical": bool(identical),
                      "values_compared": int(values), "host": res["host"], "device": res["device"],
                      "speedup_total": res["host"]["stages_ms"]["total"] / res["device"]["stages_ms"]["total"]}),
          flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--case", default="naca2412_3lvl")
    ap.add_argument("--probe-steps", type=int, default=10,
                    help="steps after the timed ones during which the hot kernels are timed in-step")
    ap.add_argument("--check-interval", type=int, default=2)
    ap.add_argument("--learned-bricks", type=int, default=2,
                    help="0 all bricks, 1 bricks with special cells and neighbours, 2 bricks with cut cells and neighbours")
    ap.add_argument("--collects", type=int, default=4)
    ap.add_argument("--pairs", type=int, default=16)
    ap.add_argument("--train-steps", type=int, default=300)
    ap.add_argument("--pretrain", type=int, default=6,
                    help="angles of attack in [0, 12] deg the network is trained on before the case (0: none)")
    ap.add_argument("--pretrain-cells-x", type=int, default=144,
                    help="base resolution of the pretraining wings")
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--smooth-wall", action="store_true",
                    help="apply the smooth-wall shear model in every step (the reference's default) and "
                         "report viscous and total loads")
    ap.add_argument("--vcycle", choices=("fp32", "bf16"), default="fp32",
                    help="V-cycle storage of the learned and the plain mixed-precision solvers")
    ap.add_argument("--cpu-budget", type=float, default=15.0, help="seconds of oracle time for the CPU baseline")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-nonorthogonal", dest="nonorthogonal", action="store_false",
                    help="skip the run with the non-orthogonal operator (flexible GMRES, FP32 V-cycle)")
    ap.add_argument("--no-autotune", dest="autotune", action="store_false",
                    help="keep the default launch shapes instead of autotuning them before the timed steps")
    ap.add_argument("--kernel-probe", type=int, default=None,
                    help="only launch hot kernel K (0 apply, 1 V-cycle sweep, 2 transport) and calibration copies")
    ap.add_argument("--host-topology", action="store_true",
                    help="build the topology on the host threads (default: device-side preprocessing)")
    ap.add_argument("--topology-build", action="store_true",
                    help="only build the case's topology on the host and on the device, check that every exported "
                         "array is identical and print the build stage times as one JSON line")
    ap.add_argument("--probe-reps", type=int, default=20)
    ap.add_argument("--probe-chunk", type=int, default=4)
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs on one GPU only; --gpus must be 1 (multi-GPU sharding is designed, "
              "not built)", file=sys.stderr)
        sys.exit(2)

    import torch
    if not torch.cuda.is_available():
        print("bench.py: no GPU visible", file=sys.stderr)
        sys.exit(1)

    from paracfd_amd import _lib as L, cases

    if args.kernel_probe is not None:
        kernel_probe(args)
        return
    if args.topology_build:
        topology_build(args)
        return
    if args.case in cases.SWEEPS:
        run_sweep(args)
        return
    g, d, spec = cases.build_case(args.case)
    t0 = time.perf_counter()
    topo = L.Topology(g, d) if args.host_topology else L.Topology(g, d, device=0)
    build_wall = time.perf_counter() - t0
    build_stages, build_device = topo.build_times()
    log(f"topology {args.case}: {topo.n_active} active cells, {topo.info.bricks} bricks, "
        f"{topo.mg_levels} MG levels, built in {build_wall:.1f}s (device {build_device})")

    vc = L.BF16 if args.vcycle == "bf16" else L.FP32
    pretrained, pretrain = pretrain_network(args, spec, vc)
    learned = run(topo, spec, args, "learned", L.FP32, vc, L.PC_LEARNED, pretrained=pretrained)
    plain32 = run(topo, spec, args, "plain", L.FP32, vc, L.PC_MG)
    plain64 = run(topo, spec, args, "plain_fp64_vcycle", L.FP32, L.FP64, L.PC_MG)
    nonorth = run(topo, spec, args, "nonorthogonal_fgmres", L.FP32, vc, L.PC_MG, nonorthogonal=1,
                  gmres_recycle=0) if args.nonorthogonal else None
    gcro = run(topo, spec, args, "nonorthogonal_gcro", L.FP32, vc, L.PC_MG, nonorthogonal=1,
               gmres_recycle=3) if args.nonorthogonal else None
    ref = run(topo, spec, args, "fp64_everywhere", L.FP64, L.FP64, L.PC_MG)

    # errors of the learned mixed-precision run against the FP64-everywhere run (same steps)
    import numpy as np
    vl = np.concatenate(learned["core"].download_cell_momentum_state())
    vr = np.concatenate(ref["core"].download_cell_momentum_state())
    pl, pr = learned["core"].download_pressure(), ref["core"].download_pressure()
    errors = {"velocity_rel_l2": rel_l2(vl, vr), "pressure_rel_l2": rel_l2(pl, pr),
              "cl_rel": abs(learned["loads"]["cl"] - ref["loads"]["cl"]) / max(abs(ref["loads"]["cl"]), 1e-30),
              "cd_rel": abs(learned["loads"]["cd"] - ref["loads"]["cd"]) / max(abs(ref["loads"]["cd"]), 1e-30)}
    log(f"errors vs FP64: {errors}")

    # roofline: in-step duration of each hot kernel on the learned solver's stream (HIP events
    # around every launch during probe steps that follow the timed ones), and the same kernels
    # launched back to back in isolation for comparison
    core = learned["core"]
    mean_its = learned["mean_its"]
    vtn = "float" if args.vcycle == "fp32" else "bf16"
    names = {5: "k_cg_dir_apply", 6: f"k_vc_update_pre2<{vtn}>", 1: f"k_smooth0<{vtn}>", 2: "k_advect<float>",
             4: "k_nn_apply"}
    probe_ids = (1, 2, 5, 6)
    torch.cuda.synchronize()
    ROCTX.push("probe:learned")
    core.probe_start(probe_ids)
    probe_its = []
    for _ in range(args.probe_steps):
        probe_its.append(core.step()["iterations"])
        core.pressure_loads()
    probed = core.probe_stop()
    torch.cuda.synchronize()
    ROCTX.pop()
    timings = {}
    for k in probe_ids + (4,):
        ms_iso, by = core.time_kernel(k, reps=20)
        launches, total = probed.get(k, (0, 0.0))
        ms = total / launches if launches else ms_iso
        per_step = launches / max(args.probe_steps, 1) if k != 4 else mean_its + 1
        timings[k] = {"ms": ms, "ms_isolated": ms_iso, "bytes": by, "launches": launches,
                      "launches_per_step": per_step, "ms_per_step": ms * per_step}
        log(f"kernel {names[k]}: in-step {ms * 1e3:.1f} us/launch over {launches} launches "
            f"({per_step:.1f}/step, {ms * per_step:.2f} ms/step), isolated {ms_iso * 1e3:.1f} us, "
            f"{by / (ms * 1e-3) / 1e9:.0f} GB/s algorithmic")
    hbm = {k: v for k, v in timings.items() if k != 4}
    dom = max(hbm, key=lambda k: hbm[k]["ms_per_step"])
    t = timings[dom]
    achieved = t["bytes"] / (t["ms"] * 1e-3) / 1e9
    roofline = {"bound": "hbm", "kernel": names[dom], "achieved": achieved, "peak": PEAK_HBM, "unit": "GB/s",
                "frac": achieved / PEAK_HBM, "traffic": measured_traffic(names[dom], args.case),
                "algorithmic_bytes_per_launch": t["bytes"], "us_per_launch_in_step": t["ms"] * 1e3,
                "launches_timed": t["launches"], "launches_per_step": t["launches_per_step"],
                "us_per_launch_isolated": t["ms_isolated"] * 1e3,
                "timing": f"HIP events around each launch on the solver's stream during {args.probe_steps} steps "
                          "after the timed ones (iterations launched one by one)",
                "traffic_source": "profiles/traffic.json: rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE of this kernel on this "
                                  "workload, scaled by the counters of a device copy of known size"}
    # whole step: algorithmic bytes per active cell of one step (each pass touching its arrays
    # once; FP32 fields, FP32 V-cycle) over the measured step time
    subs = learned["substeps"]
    n_patch = topo.info.patches
    # the reference algorithm's minimum bytes per pass (this build's passes move fewer in places:
    # Fletcher-Reeves needs no r_prev, z is stored in FP32 where the V-cycle computes it so)
    b_cell = (18 + 43 + 24 * subs + 26 + 30 + mean_its * (98 + 114) + 34 + 38 + 16 + 80.0 * n_patch / topo.n_active)
    step_gbs = b_cell * topo.n_active / (learned["ms_per_step"] * 1e-3) / 1e9
    step_roofline = {"bytes_per_cell": b_cell, "achieved": step_gbs, "peak": PEAK_HBM, "unit": "GB/s",
                     "frac": step_gbs / PEAK_HBM,
                     "model": "18 CFL + 43 transport + 24 per diffusion substep + 26 reconstruction + 30 RHS + "
                              "iterations x (98 PCG + 114 FP32 V-cycle) + 34 correction + 38 impulse + 16 divergence "
                              "check + 80 per surface patch, bytes per active cell"}
    log(f"step bytes model: {b_cell:.0f} B/cell -> {step_gbs:.0f} GB/s over the measured step ({step_gbs / PEAK_HBM:.3f} of peak)")
    ceilings = device_ceilings()
    log(f"device ceilings: copy {ceilings['copy_GBps']:.0f} GB/s, BF16 GEMM {ceilings['bf16_gemm_TFLOPs']:.0f} TFLOP/s")
    roofline["frac_of_measured_copy"] = achieved / ceilings["copy_GBps"]
    tr = learned["train"]
    ms_nn = timings[4]["ms"]
    # useful network flops per launch: layers 0-2 16 -> 16 channels, layer 3 16 -> 1 (channel 0)
    flops_nn = tr["bricks"] * 512 * 27 * 16 * (3 * 16 + 1) * 2.0
    nn_roof = {"bound": "mfma", "kernel": "k_nn_apply", "achieved": flops_nn / (ms_nn * 1e-3) / 1e12,
               "peak": PEAK_BF16, "unit": "TFLOP/s", "frac": flops_nn / (ms_nn * 1e-3) / 1e12 / PEAK_BF16,
               "us_per_launch": ms_nn * 1e3, "launches_per_step": timings[4]["launches_per_step"]}
    nn_roof["frac_of_measured_gemm"] = nn_roof["achieved"] / ceilings["bf16_gemm_TFLOPs"]

    cpu = None
    if not args.no_cpu_baseline:
        from paracfd_amd import default_options
        cpu = cpu_baseline(topo, default_options(rho=spec["rho"], nu=spec["nu"], speed=spec["speed"],
                                                 cfl=spec["cfl"], rtol=1e-8, max_iterations=500,
                                                 smooth_wall=1 if args.smooth_wall else 0), args.cpu_budget)
        log(f"cpu baseline {cpu['value']:.4g} cell-steps/s")

    def summary(r):
        return {"value": r["value"], "ms_per_step": r["ms_per_step"], "mean_pcg_iterations": r["mean_its"],
                "converged": r["converged"], "cl": r["loads"]["cl"], "cd": r["loads"]["cd"]}

    n = topo.n_active
    out = {
        "metric": METRIC, "value": learned["value"], "unit": "cell-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": learned["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": f"f32 fields + {args.vcycle} V-cycle + bf16 network + f64 PCG",
        "data": "synthetic (analytic geometry, freestream start)",
        "config": {"workload": args.case, "active_cells": n, "amr_levels": topo.info.levels,
                   "preconditioner": "V-cycle + learned 4-layer 16-channel conv correction (weights per AMR level, neighbour-brick input halo), flexible PCG",
                   "network_bricks": tr["bricks"], "rtol": 1e-8, "mean_pcg_iterations": mean_its,
                   "fallback_steps": learned["fallback_steps"], "network_steps": learned["learned_steps"],
                   "network_rejected_by_validation": learned["learned_rejected"], "cl": learned["loads"]["cl"],
                   "cd": learned["loads"]["cd"], "smooth_wall": bool(args.smooth_wall),
                   "cd_viscous": learned["loads"]["cd_viscous"], "total_cd": learned["loads"]["total_cd"],
                   "total_cl": learned["loads"]["total_cl"],
                   "training": {k: tr[k] for k in ("pairs", "train_steps", "loss_first", "loss_last", "train_ms")},
                   "pretraining": pretrain,
                   "launch_shapes": learned["tuning"],
                   "plain_fp32_vcycle": summary(plain32), "plain_fp64_vcycle": summary(plain64),
                   "nonorthogonal_fgmres": summary(nonorth) if nonorth else None,
                   "nonorthogonal_gcro_recycle3": summary(gcro) if gcro else None,
                   "speedup_vs_plain_fp32_vcycle": learned["value"] / plain32["value"],
                   "speedup_vs_plain_fp64_vcycle": learned["value"] / plain64["value"],
                   "errors_vs_fp64": errors},
        "topology_build": {"device": build_device, "wall_s": build_wall,
                           "stages_ms": {k: round(v, 1) for k, v in build_stages.items()}},
        "roofline": roofline,
        "step_roofline": step_roofline,
        "roofline_network": nn_roof,
        "device_ceilings": ceilings,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
