"""Benchmark of the per-time-step AMR D3Q19 update on one MI355X.

Workload (default): the 2-level cylinder channel -- coarse 128x32x32 cells
in 8^3 blocks, one refinement level over x in [xc-2D, xc+6D], z-spanning
cylinder of D = 16 fine cells with Bouzidi links, velocity inlet u = 0.04,
pressure outlet, tau = 0.55.  A step is one coarse time step
(TimeLoop::advance_one: level 0 once, level 1 twice, coupling, FD boundaries,
drag snapshot).  Inputs are synthetic (built in memory), resident in HBM
before the timed region; the timed region replays hipGraphs of the step.

Prints one JSON line: FP32 fine-equivalent MLUPS as `value`, the FP64 figure
and the FP32/FP64 ratio beside it, the roofline of the fused collide+stream
kernel (HIP-event timing on its stream) and a CPU baseline (the FP64 oracle
restatement on the host, bounded sample).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
BYTES_PER_UPDATE = {64: 304.0, 32: 152.0}  # 19 populations read + 19 written


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def build_case(name):
    from octlb_amd import cases

    if name == "cylinder":
        return cases.cylinder_channel(bricks_x=16, bricks_yz=4, base_level=0, fine=True, xc=1.0,
                                      radius=0.125, u_in=0.04, tau=0.55)
    if name == "cavity":
        return cases.cavity(64, tau=0.6, u_lid=0.05)
    raise SystemExit(f"unknown workload {name}")


def run_gpu(case, precision, steps, warmup, timing_steps):
    from octlb_amd.solver import GpuSolver
    from oracle.binding import setup  # applies kinds / links / init (shared with tests)

    sim = GpuSolver(case.topo, case.params(precision))
    setup(sim, case)
    actual, feq = sim.update_counts()
    sim.advance_graph(max(warmup, 1))
    sim.synchronize()
    t0 = time.perf_counter()
    sim.advance_graph(steps)
    sim.synchronize()
    dt = time.perf_counter() - t0
    # per-launch duration of the fused kernel: eager launches, HIP events on
    # the solver's stream
    sim.enable_timing(True)
    sim.advance(timing_steps)
    ms, launches, updates = sim.kernel_timing()
    sim.enable_timing(False)
    drag = sim.drag(case.drag_bbox) if case.drag_bbox is not None else np.zeros(3)
    mass = sim.mass()
    res = {
        "seconds": dt,
        "ms_per_step": 1e3 * dt / steps,
        "mlups_feq": feq * steps / dt / 1e6,
        "mlups_actual": actual * steps / dt / 1e6,
        "kernel_ms": ms,
        "kernel_launches": launches,
        "kernel_updates": updates,
        "drag": drag.tolist(),
        "mass": mass,
        "actual_per_step": actual,
        "feq_per_step": feq,
    }
    sim.close()
    return res


def cpu_baseline(case, budget_s):
    """FP64 oracle restatement, single thread, on a bounded number of steps."""
    from oracle.binding import OracleSim, setup

    sim = OracleSim(case.topo, case.params(64))
    setup(sim, case)
    sim.set_threads(1)
    nb = case.topo.num_blocks
    lv = case.topo.level
    actual = float(np.sum(case.n ** 3 * 2.0 ** lv))
    feq = float(np.sum(case.n ** 3 * 8.0 ** (lv.max() - lv)) * 2.0 ** lv.max())
    t0 = time.perf_counter()
    sim.advance(1)
    one = time.perf_counter() - t0
    steps = max(1, min(200, int(budget_s / max(one, 1e-6))))
    t0 = time.perf_counter()
    sim.advance(steps)
    dt = time.perf_counter() - t0
    sim.close()
    return {"value": feq * steps / dt / 1e6, "unit": "MLUPS (fine-equivalent)", "cores": 1,
            "kind": "port", "actual_mlups": actual * steps / dt / 1e6,
            "sample": f"{steps} coarse steps of the same workload ({nb} blocks), FP64 oracle "
                      f"restatement, 1 thread, {dt:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="cylinder")
    ap.add_argument("--timing-steps", type=int, default=50)
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--precision", default="both", choices=["both", "32", "64"])
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is a design "
                                   "note only (DESIGN.md), no figure at N > 1 is produced",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    log("building case")
    case = build_case(args.workload)
    res = {}
    precs = [32, 64] if args.precision == "both" else [int(args.precision)]
    for p in precs:
        log(f"gpu fp{p}: warmup {args.warmup}, timed {args.steps} coarse steps")
        res[p] = run_gpu(case, p, args.steps, args.warmup, args.timing_steps)
        log(f"  fp{p}: {res[p]['mlups_feq']:.1f} MLUPS feq, {res[p]['ms_per_step']:.4f} ms/step")
    head = 32 if 32 in res else precs[0]
    r = res[head]
    kb = BYTES_PER_UPDATE[head] * r["kernel_updates"]
    achieved = kb / (r["kernel_ms"] * 1e-3) / 1e9 if r["kernel_ms"] > 0 else 0.0
    cpu = None
    if not args.no_cpu:
        log("cpu baseline")
        cpu = cpu_baseline(case, args.cpu_budget)
    out = {
        "metric": "FP32 vs FP64 MLUPS & % HBM roofline, 1 MI355X; adjoint/forward cost ratio",
        "value": round(r["mlups_feq"], 3),
        "unit": "MLUPS (fine-equivalent)",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["ms_per_step"], 6),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if head == 32 else "f64",
        "data": "synthetic",
        "config": {
            "workload": ("cylinder channel, 2-level AMR, coarse 128x32x32 in 8^3 blocks, fine box "
                         "2D upstream to 6D downstream, D=16 fine cells, Bouzidi links, inlet "
                         "u=0.04, pressure outlet, tau=0.55")
            if args.workload == "cylinder" else "lid-driven cavity 64^3, tau=0.6, lid 0.05",
            "blocks": case.topo.num_blocks,
            "levels": int(case.topo.max_level) + 1,
            "actual_updates_per_step": r["actual_per_step"],
            "fine_equivalent_updates_per_step": r["feq_per_step"],
            "timed_region": "hipGraph replay of whole coarse steps",
        },
        "mlups_actual": round(r["mlups_actual"], 3),
        "fp32_mlups": round(res[32]["mlups_feq"], 3) if 32 in res else None,
        "fp64_mlups": round(res[64]["mlups_feq"], 3) if 64 in res else None,
        "fp32_over_fp64": (round(res[32]["mlups_feq"] / res[64]["mlups_feq"], 3)
                           if 32 in res and 64 in res else None),
        "adjoint_over_forward": None,
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None,
            "kernel": "fused collide+stream (KBlock)",
            "launches": r["kernel_launches"],
            "avg_launch_us": round(1e3 * r["kernel_ms"] / max(r["kernel_launches"], 1), 3),
            "bytes_per_update": BYTES_PER_UPDATE[head],
        },
        "cpu_baseline": cpu,
        "drag_fp32": res[32]["drag"] if 32 in res else None,
        "drag_fp64": res[64]["drag"] if 64 in res else None,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
