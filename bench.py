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
    if name == "synthetic":
        return cases.synthetic_amr(bricks=8, max_level=2, frac=(0.5, 0.25), seed=7)
    raise SystemExit(f"unknown workload {name}")


def traffic_bytes(workload, precision):
    """HBM bytes per launch of the dominant kernel, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE passes (profiles/kblock_traffic.json,
    which records the commands and the copy-kernel scale factors), or None."""
    path = ROOT / "profiles" / "kblock_traffic.json"
    if not path.exists():
        return None
    rec = json.loads(path.read_text()).get(f"{workload}/fp{precision}")
    if not rec:
        return None
    kib = rec["fetch_kib"] * rec["fetch_scale"] + rec["write_kib"] * rec["write_scale"]
    return round(kib * 1024.0)


SCHEMES = {"auto": None, "pull": 0, "push": 1}


def choose_scheme(sim, scheme, reps):
    """The fused kernel's update scheme: the autotuner's choice, or the one
    asked for (the autotuner's timings are then not taken)."""
    if SCHEMES[scheme] is None:
        return sim.autotune(reps=reps)
    sim.set_kernel_variant(SCHEMES[scheme])
    return SCHEMES[scheme], None


def epilogue_record(sim):
    """Which coupling steps run inside the update launches (bit 0: half-time
    prolongation, bit 1: restriction) and the autotuner's weighted times."""
    mode, ms = sim.coupling_epilogues()
    return {"half_time_prolongation": bool(mode & 1), "restriction": bool(mode & 2),
            "ms_per_coarse_step": {"separate": round(float(ms[0]), 4), "chosen": round(float(ms[1]), 4)}}


def run_gpu(case, precision, steps, warmup, timing_steps, scheme="auto"):
    from octlb_amd.solver import GpuSolver

    sim = GpuSolver(case.topo, case.params(precision))
    case.apply(sim)
    actual, feq = sim.update_counts()
    variant, tuned_ms = choose_scheme(sim, scheme, 10)  # update scheme of the fused kernel, state unchanged
    sim.prepare_graphs()  # every graph the timed region replays exists before the timer
    sim.advance_graph(max(warmup, 1))
    sim.synchronize()
    t0 = time.perf_counter()
    sim.advance_graph(steps)
    sim.synchronize()
    dt = time.perf_counter() - t0
    drag = sim.drag(case.drag_bbox) if case.drag_bbox is not None else np.zeros(3)
    drag_us = None
    if case.drag_bbox is not None:
        # one drag query as a caller sees it: kernels, 3 doubles back, host sync
        t0 = time.perf_counter()
        for _ in range(20):
            sim.drag(case.drag_bbox)
        drag_us = 1e6 * (time.perf_counter() - t0) / 20
    mass = sim.mass()
    # per-launch duration of the fused kernel of every level: back-to-back
    # replays between HIP events on the solver's stream (state unchanged)
    levels = {}
    for lv in sorted(set(int(x) for x in case.topo.level)):
        ms, cells = sim.time_kernel(lv, timing_steps)
        levels[lv] = {"avg_launch_us": 1e3 * ms, "cells_per_launch": cells}
    res = {
        "seconds": dt,
        "ms_per_step": 1e3 * dt / steps,
        "mlups_feq": feq * steps / dt / 1e6,
        "mlups_actual": actual * steps / dt / 1e6,
        "kernel_levels": levels,
        "drag": drag.tolist(),
        "drag_query_us": drag_us,
        "mass": mass,
        "actual_per_step": actual,
        "feq_per_step": feq,
        "autotune": {"variant": {0: "pull", 1: "push where plain"}[variant], "push_blocks": sim.kernel_variant()[1],
                     "fused_ms_per_coarse_step": None if tuned_ms is None else
                     {"pull": round(float(tuned_ms[0]), 4), "push where plain": round(float(tuned_ms[1]), 4)},
                     "chosen_by": "fixed" if tuned_ms is None else "autotuner",
                     "coupling_epilogues": epilogue_record(sim)},
    }
    sim.close()
    return res


def run_ensemble(precs, steps, warmup, timing_steps, count=64, scheme="auto"):
    """64 independent lid-driven cavities (cells per side drawn from
    {32, 48, 64, 96}, tau uniform in [0.52, 0.8], lid speed uniform in
    [0.02, 0.08], one seed) stepped as one device state: one fused-kernel
    launch, one FD launch and the per-member ConstRho averages per step,
    replayed as hipGraphs."""
    from octlb_amd import cases
    from octlb_amd.solver import GpuSolver

    ens = cases.cavity_ensemble(count=count)
    cells = int(ens.topo.num_blocks) * 512
    out = {"workload": ("64 lid-driven cavities in one launch, cells per side from {32,48,64,96}, tau "
                        "U[0.52,0.8], lid U[0.02,0.08], seed 11, FD walls, ConstRhoBGK with per-member "
                        "tree-reduced averages"),
           "members": count, "blocks": int(ens.topo.num_blocks), "cells": cells,
           "sizes": [int(c.info["cells"]) for c in ens.members]}
    for p in precs:
        sim = GpuSolver.ensemble(ens, p)
        sim.set_bc_kinds(ens.kinds)
        sim.init_equilibrium(1.0, (0.0, 0.0, 0.0), 0xFF)
        variant, tuned_ms = choose_scheme(sim, scheme, 5)
        sim.prepare_graphs()
        sim.advance_graph(max(warmup, 1))
        sim.synchronize()
        t0 = time.perf_counter()
        sim.advance_graph(steps)
        sim.synchronize()
        dt = time.perf_counter() - t0
        ms, upd = sim.time_kernel(0, timing_steps)
        ach = BYTES_PER_UPDATE[p] * upd / (ms * 1e-3) / 1e9
        out[f"fp{p}"] = {"mlups": round(cells * steps / dt / 1e6, 3), "ms_per_step": round(1e3 * dt / steps, 6),
                         "steps": steps,
                         "autotune": {"variant": {0: "pull", 1: "push where plain"}[variant],
                                      "fused_ms": None if tuned_ms is None else [round(float(x), 4) for x in tuned_ms],
                                      "chosen_by": "fixed" if tuned_ms is None else "autotuner"},
                         "roofline": {"bound": "hbm", "achieved": round(ach, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                      "frac": round(ach / HBM_PEAK_GBS, 4), "kernel": "fused collide+stream (KBlock)",
                                      "cells_per_launch": upd, "avg_launch_us": round(1e3 * ms, 3)}}
        sim.close()
    if 32 in precs and 64 in precs:
        out["fp32_over_fp64"] = round(out["fp32"]["mlups"] / out["fp64"]["mlups"], 3)
    return out


def run_adjoint_channel(steps, checkpoint):
    """Discrete adjoint of the benchmarked 2-level cylinder channel itself
    (octlb_adj_run, multi-level engine: Lagrava coupler, FD velocity inlet
    and pressure outlet, Bouzidi links).  J = mean x-drag over the second
    half of `steps` coarse steps from rest; gradients w.r.t. the inlet
    velocity and every link q; checkpoint every `checkpoint` coarse steps.
    The ratio divides the adjoint run's device time by the same number of
    production FP64 coarse steps (hipGraph replay).  The FP32-forward mode
    (FP32 states, FP64 adjoint) is reported beside it with its gradient
    error against the FP64 mode."""
    from octlb_amd.adjoint import DragAdjoint
    from octlb_amd.solver import GpuSolver

    case = build_case("cylinder")
    win = (steps // 2, steps)
    adj = DragAdjoint(case, steps, win, checkpoint=checkpoint)
    J, _, dq = adj.evaluate()
    fw, bw = adj.timing
    gi = adj.inlet_grad.copy()
    adj32 = DragAdjoint(case, steps, win, checkpoint=checkpoint, precision=32)
    J32, _, dq32 = adj32.evaluate()
    fw32, bw32 = adj32.timing
    gi32 = adj32.inlet_grad.copy()
    sim = case.apply(GpuSolver(case.topo, case.params(64)))
    sim.autotune(reps=5)
    sim.prepare_graphs()
    sim.advance_graph(16)
    sim.synchronize()
    t0 = time.perf_counter()
    sim.advance_graph(steps)
    sim.synchronize()
    fwd_ms = 1e3 * (time.perf_counter() - t0) / steps
    sim.close()
    return {"workload": "the benchmarked 2-level cylinder channel (1,152 blocks), from rest, inlet u = 0.04",
            "coarse_steps": steps, "window": list(win), "checkpoint": checkpoint, "links": int(len(case.q)),
            "J": J, "dJ_du_inlet": [float(v) for v in gi], "max_abs_dJ_dq": float(np.abs(dq).max()),
            "adjoint_ms": round(fw + bw, 3), "forward_sweep_ms": round(fw, 3), "reverse_sweep_ms": round(bw, 3),
            "tapes": "one taping pass per segment" if adj.packed else "per-step re-run",
            "forward_fp64_ms_per_coarse_step": round(fwd_ms, 5),
            "adjoint_over_forward_fp64": round((fw + bw) / (steps * fwd_ms), 3),
            "fp32_forward": {"adjoint_ms": round(fw32 + bw32, 3),
                             "over_forward_fp64": round((fw32 + bw32) / (steps * fwd_ms), 3),
                             "rel_err_vs_fp64_adjoint": {
                                 "J": abs(J32 - J) / abs(J),
                                 "dJ_du_inlet": abs(gi32[0] - gi[0]) / abs(gi[0]),
                                 "dJ_dq_max": float(np.abs(dq32 - dq).max() / np.abs(dq).max())}}}


def adjoint_ratio(adjc, tol=1e-4):
    """The adjoint-to-forward cost ratio of the channel run, against the
    production FP64 coarse step: the FP32-forward mode (FP32 states, FP64
    adjoint) when its J and gradients stay within tol of the FP64 adjoint's,
    else the FP64 mode."""
    if not adjc:
        return None
    f32 = adjc["fp32_forward"]
    ok = all(v <= tol for v in f32["rel_err_vs_fp64_adjoint"].values())
    return {"value": f32["over_forward_fp64"] if ok else adjc["adjoint_over_forward_fp64"],
            "mode": "FP32 forward states, FP64 adjoint" if ok else "FP64 states and adjoint",
            "fp64_mode": adjc["adjoint_over_forward_fp64"],
            "fp32_forward_mode": f32["over_forward_fp64"],
            "fp32_forward_max_rel_err": max(f32["rel_err_vs_fp64_adjoint"].values()),
            "against": "production FP64 coarse step (hipGraph replay) of the same channel"}


def run_adjoint(steps, checkpoint):
    """Discrete adjoint of the single-level step (octlb_adj_run) on the
    cylinder at the fine resolution of the 2-level channel, as one level:
    256 x 64 x 64 cells in 2,048 blocks of 8^3, D = 16 cells, Bouzidi on all
    cut links, bounce-back walls (closed box, fluid started at
    feq(1, (0.04, 0, 0))), tau 0.55.  J = mean x-drag over the second half of
    the run; gradients w.r.t. u0 and every link q, checkpoints every
    `checkpoint` steps.  A step is one update of the blocks' level (level 3
    here, so a coarse step of octlb_dev_advance is 8 of them).  The ratio
    divides the adjoint run's device time (forward sweep + reverse sweep with
    recompute) by the same number of production FP64 updates of the same case
    (hipGraph replay of coarse steps)."""
    from octlb_amd import cases
    from octlb_amd.adjoint import DragAdjoint
    from octlb_amd.solver import GpuSolver

    case = cases.cylinder_channel(bricks_x=4, base_level=3, fine=False, bricks_yz=1, xc=1.0, radius=0.125,
                                  u_in=0.04, tau=0.55, open_ends=False)
    win = (steps // 2, steps)
    adj = DragAdjoint(case, steps, win, checkpoint=checkpoint)
    adj.evaluate()
    t0 = time.perf_counter()
    J, du, dq = adj.evaluate()
    wall = time.perf_counter() - t0
    fw, bw = adj.timing
    # FP32 forward states, FP64 adjoint (the C5 precision mode)
    adj32 = DragAdjoint(case, steps, win, checkpoint=checkpoint, precision=32)
    J32, du32, dq32 = adj32.evaluate()
    fw32, bw32 = adj32.timing
    per_coarse = 2 ** int(case.topo.level.max())  # updates of the one level per coarse step
    coarse = max(1, steps // per_coarse)
    fwd = {}
    for p in (64, 32):
        sim = GpuSolver(case.topo, case.params(p))
        sim.set_bc_kinds(case.kinds)
        sim.set_links(case.link, case.q)
        sim.init_equilibrium(1.0, case.init_u0, case.init_mask)
        sim.prepare_graphs()
        sim.advance_graph(4)
        sim.synchronize()
        t0 = time.perf_counter()
        sim.advance_graph(coarse)
        sim.synchronize()
        fwd[p] = 1e3 * (time.perf_counter() - t0) / (coarse * per_coarse)
        sim.close()
    return {"workload": ("cylinder D=16 cells in a closed 256x64x64 box (2,048 blocks of 8^3, one level), "
                         "Bouzidi on all cut links, bounce-back walls, u0 = 0.04, tau 0.55"),
            "steps": steps, "window": list(win), "checkpoint": checkpoint, "links": int(len(case.q)),
            "J": J, "dJ_du0": [float(v) for v in du], "max_abs_dJ_dq": float(np.abs(dq).max()),
            "adjoint_ms": round(fw + bw, 3), "forward_sweep_ms": round(fw, 3), "reverse_sweep_ms": round(bw, 3),
            "adjoint_wall_ms": round(1e3 * wall, 3),
            "forward_fp64_ms_per_update": round(fwd[64], 5), "forward_fp32_ms_per_update": round(fwd[32], 5),
            "adjoint_over_forward_fp64": round((fw + bw) / (steps * fwd[64]), 3),
            "adjoint_over_forward_fp32": round((fw + bw) / (steps * fwd[32]), 3),
            "fp32_forward": {"adjoint_ms": round(fw32 + bw32, 3), "forward_sweep_ms": round(fw32, 3),
                             "reverse_sweep_ms": round(bw32, 3),
                             "over_forward_fp64": round((fw32 + bw32) / (steps * fwd[64]), 3),
                             "over_forward_fp32": round((fw32 + bw32) / (steps * fwd[32]), 3),
                             "rel_err_vs_fp64_adjoint": {
                                 "J": abs(J32 - J) / abs(J),
                                 "dJ_du0": abs(du32[0] - du[0]) / abs(du[0]),
                                 "dJ_dq_max": float(np.abs(dq32 - dq).max() / np.abs(dq).max())}}}


def cpu_model():
    """The host CPU's model string (lscpu's 'Model name', from /proc/cpuinfo)."""
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_threads():
    """Host threads this process may use: the OpenMP thread count the
    environment sets, else the CPUs the process may run on."""
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        return int(env)
    return len(os.sched_getaffinity(0))


def cpu_leg(case, budget_s, threads):
    """FP64 oracle restatement on `threads` host threads (OpenMP over blocks),
    timed on a bounded number of coarse steps of the same workload."""
    from oracle.binding import OracleSim

    sim = case.apply(OracleSim(case.topo, case.params(64)))
    sim.set_threads(threads)
    lv = case.topo.level
    actual = float(np.sum(case.n ** 3 * 2.0 ** lv))
    feq = float(np.sum(case.n ** 3 * 8.0 ** (lv.max() - lv)) * 2.0 ** lv.max())
    t0 = time.perf_counter()
    sim.advance(1)
    one = time.perf_counter() - t0
    steps = max(1, min(400, int(budget_s / max(one, 1e-6))))
    t0 = time.perf_counter()
    sim.advance(steps)
    dt = time.perf_counter() - t0
    sim.close()
    return steps, dt, feq * steps / dt / 1e6, actual * steps / dt / 1e6


def cpu_baseline(case, budget_s):
    """FP64 oracle restatement on the host: one thread (the reference is
    single-threaded per MPI rank) and OpenMP over blocks on every host thread
    this process may use; the OpenMP figure is `value`."""
    nb = case.topo.num_blocks
    threads = cpu_threads()
    s1, dt1, feq1, act1 = cpu_leg(case, budget_s, 1)
    sn, dtn, feqn, actn = cpu_leg(case, budget_s, threads)
    return {"value": feqn, "unit": "MLUPS (fine-equivalent)", "cores": threads,
            "kind": "port", "actual_mlups": actn,
            "sample": f"{sn} coarse steps of the same workload ({nb} blocks), FP64 oracle "
                      f"restatement, OpenMP over blocks on {threads} threads, {dtn:.1f} s",
            "cpu_model": cpu_model(), "nproc": os.cpu_count(),
            "single_thread": {"value": feq1, "actual_mlups": act1, "cores": 1,
                              "sample": f"{s1} coarse steps, 1 thread, {dt1:.1f} s"}}


FP64_VALU_PEAK_GFLOPS = 78600.0
FLOPS_PER_CELL_TRIANGLE = 101  # 3 edges x (closest point 23 + distance 8) + distance 8


def run_migrate(precision=32, mark_frac=0.1, seed=3, reps=3):
    """State migration after a forest change (octlb_dev_migrate) on the
    synthetic 3-level forest: a seeded 10% of its level-1 blocks refined
    once more and 2:1 balanced, the state moved onto it (copies +
    refinements) and back (copies + coarsenings).  Wall time per migration,
    host block matching, task upload and the device launch included; the
    rate counts each old block's populations read once and each new
    block's written once."""
    from octlb_amd import cases
    from octlb_amd.solver import GpuSolver

    case = build_case("synthetic")
    f = cases.synthetic_forest(bricks=8, max_level=2, frac=(0.5, 0.25), seed=7)
    ta = f.topology(8, 1)
    assert ta.num_blocks == case.topo.num_blocks
    rng = np.random.default_rng(seed)
    mark = ((ta.level == 1) & (rng.random(ta.num_blocks) < mark_frac)).astype(np.uint8)
    f.refine_marked(mark)
    f.balance()
    tb = f.topology(8, 1)
    src = case.apply(GpuSolver(case.topo, case.params(precision)))
    src.advance(2)
    dst = GpuSolver(tb, case.params(precision))
    dst.set_bc_kinds(cases.walls_kinds(tb))
    back = GpuSolver(case.topo, case.params(precision))
    back.set_bc_kinds(case.kinds)
    dst.migrate_from(src)
    back.migrate_from(dst)
    fw, bw = [], []
    for _ in range(reps):
        t0 = time.perf_counter()
        dst.migrate_from(src)
        t1 = time.perf_counter()
        back.migrate_from(dst)
        fw.append(t1 - t0)
        bw.append(time.perf_counter() - t1)
    for sim in (src, dst, back):
        sim.close()
    elem = 4 if precision == 32 else 8
    per_block = 19 * 512 * elem
    na, nb_ = ta.num_blocks, tb.num_blocks
    fms, bms = 1e3 * min(fw), 1e3 * min(bw)
    return {"workload": ("synthetic 3-level forest (8^3 bricks) -> 10% of its level-1 blocks refined "
                         "once more, 2:1 balanced -> back"),
            "precision": f"fp{precision}", "blocks": [na, nb_],
            "marked_level1_blocks": int(mark.sum()),
            "forward_ms": round(fms, 3), "back_ms": round(bms, 3),
            "forward_gbs": round((na + nb_) * per_block / (fms * 1e-3) / 1e9, 1),
            "back_gbs": round((na + nb_) * per_block / (bms * 1e-3) / 1e9, 1)}


def run_link_build(with_cpu, segments=512, rings=8, reps=3, cpu_blocks=24):
    """The geometry pipeline on the headline forest: a z-spanning cylinder of
    D = 16 fine cells as 2 * segments * rings triangles, voxelized into
    MaterialKind per cell (voxelize_part, octlb_voxelize_part) and turned
    into Bouzidi links (BouzidiLinkData::Build, octlb_build_links).  GPU:
    wall time of each C call (uploads, kernels, downloads).  CPU: the oracle
    restatements -- the voxelizer on all host threads over the whole
    forest, the link build (the reference's loop: every fluid cell against
    every triangle) on `cpu_blocks` blocks, scaled per fluid cell."""
    from octlb_amd import cases, geometry

    case = cases.cylinder_channel(bricks_x=16, bricks_yz=4, base_level=0, fine=True, xc=1.0,
                                  radius=0.125, u_in=0.04, tau=0.55)
    topo = case.topo
    tris = geometry.cylinder_soup(1.0, 0.5, 0.125, 0.0, 1.0, segments, rings)
    mat, nexact = geometry.voxelize_part(topo.bounds, 8, tris, geometry.PART_EXTERNAL_OBSTACLE)  # warm-up
    t0 = time.perf_counter()
    for _ in range(reps):
        mat, nexact = geometry.voxelize_part(topo.bounds, 8, tris, geometry.PART_EXTERNAL_OBSTACLE)
    vox_s = (time.perf_counter() - t0) / reps
    vox = {"gpu_ms": round(vox_s * 1e3, 3), "cells": int(mat.size),
           "fluid_solid_boundary": [int((mat == k).sum()) for k in range(3)],
           "exact_fallback_cells": nexact}
    if with_cpu:
        from oracle.binding import voxelize_part as oracle_voxelize

        t0 = time.perf_counter()
        omat = oracle_voxelize(topo.bounds, 8, tris, geometry.PART_EXTERNAL_OBSTACLE)
        vox["cpu_baseline"] = {"value_s": round(time.perf_counter() - t0, 3), "cores": cpu_threads(), "kind": "port",
                               "sample": "oracle restatement (exact arithmetic), whole forest, OpenMP over blocks"}
        vox["equal_to_oracle"] = bool(np.array_equal(omat, mat))
    link, q = geometry.build_links(topo.bounds, 8, mat, tris)  # warm-up (code object, allocator)
    t0 = time.perf_counter()
    for _ in range(reps):
        link, q = geometry.build_links(topo.bounds, 8, mat, tris)
    gpu_s = (time.perf_counter() - t0) / reps
    # candidate cells: fluid with a solid-like in-block neighbour
    solid = mat != 0
    pad = np.pad(solid, ((0, 0), (1, 1), (1, 1), (1, 1)), constant_values=False)
    near = np.zeros_like(solid)
    for i in range(1, 19):
        ex, ey, ez = cases.C[i]
        near |= pad[:, 1 + ez:9 + ez, 1 + ey:9 + ey, 1 + ex:9 + ex]
    ncand = int((near & ~solid).sum())
    gflops = ncand * tris.shape[0] * FLOPS_PER_CELL_TRIANGLE / gpu_s / 1e9
    out = {
        "workload": ("cylinder-channel forest (1,152 blocks of 8^3), z-spanning cylinder D=16 fine cells "
                     f"as {tris.shape[0]} triangles ({segments} segments x {rings} rings); material voxelized "
                     "from the soup on the device, links from that material"),
        "triangles": int(tris.shape[0]),
        "candidate_cells": ncand,
        "links": int(len(link)),
        "gpu_ms": round(gpu_s * 1e3, 3),
        "voxelize": vox,
        "roofline": {"bound": "fp64 valu", "achieved": round(gflops, 1), "peak": FP64_VALU_PEAK_GFLOPS,
                     "unit": "GFLOP/s", "frac": round(gflops / FP64_VALU_PEAK_GFLOPS, 4),
                     "flops_per_cell_triangle": FLOPS_PER_CELL_TRIANGLE},
    }
    if with_cpu:
        from oracle.binding import build_links as oracle_links

        fluid = (~solid).reshape(topo.num_blocks, -1).sum(1)
        order = np.argsort(-near.reshape(topo.num_blocks, -1).sum(1), kind="stable")[:cpu_blocks]
        t0 = time.perf_counter()
        oracle_links(topo.bounds[order], 8, mat[order], tris)
        dt = time.perf_counter() - t0
        per_cell = dt / max(int(fluid[order].sum()), 1)
        out["cpu_baseline"] = {"est_full_s": round(per_cell * int(fluid.sum()), 2), "cores": 1, "kind": "port",
                               "sample": (f"oracle restatement on the {cpu_blocks} blocks with most candidate "
                                          f"cells ({int(fluid[order].sum())} fluid cells, {dt:.1f} s), "
                                          "scaled per fluid cell to the whole forest")}
        out["speedup_vs_cpu"] = round(out["cpu_baseline"]["est_full_s"] / gpu_s, 1)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="cylinder")
    ap.add_argument("--timing-steps", type=int, default=50,
                    help="back-to-back replays of each level's fused kernel for its launch time")
    ap.add_argument("--cpu-budget", type=float, default=10.0,
                    help="seconds of host work per CPU-baseline leg (1 thread, all threads)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-synthetic", action="store_true",
                    help="skip the 3-level synthetic AMR measurement reported beside the headline")
    ap.add_argument("--synthetic-steps", type=int, default=40)
    ap.add_argument("--no-links", action="store_true",
                    help="skip the Bouzidi link-build measurement")
    ap.add_argument("--no-migrate", action="store_true",
                    help="skip the state-migration (forest change) measurement")
    ap.add_argument("--no-ensemble", action="store_true",
                    help="skip the 64-cavity ensemble measurement")
    ap.add_argument("--ensemble-steps", type=int, default=40)
    ap.add_argument("--no-adjoint", action="store_true",
                    help="skip the adjoint/forward cost measurement")
    ap.add_argument("--adjoint-steps", type=int, default=4000,
                    help="adjoint run length; J averages the drag over its second half")
    ap.add_argument("--adjoint-checkpoint", type=int, default=100)
    ap.add_argument("--channel-adjoint-steps", type=int, default=4000,
                    help="coarse steps of the 2-level channel's adjoint run (J over the second half)")
    ap.add_argument("--precision", default="both", choices=["both", "32", "64"])
    ap.add_argument("--scheme", default="auto", choices=sorted(SCHEMES),
                    help="fused-kernel update scheme: the autotuner's choice (default) or a fixed one")
    ap.add_argument("--calibrate-copy-mb", type=int, default=0,
                    help="run one device-to-device copy of this many MiB first (profiler "
                         "counter scale: a plain copy reads and writes exactly that many bytes)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; sharding over several GPUs is a design "
                                   "note only, no figure at N > 1 is produced",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    if args.calibrate_copy_mb:
        import torch

        nbytes = args.calibrate_copy_mb << 20
        src = torch.empty(nbytes, dtype=torch.uint8, device="cuda").fill_(1)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"calibration copy: {nbytes} bytes")
        del src, dst
    log("building case")
    case = build_case(args.workload)
    res = {}
    precs = [32, 64] if args.precision == "both" else [int(args.precision)]
    for p in precs:
        log(f"gpu fp{p}: warmup {args.warmup}, timed {args.steps} coarse steps")
        res[p] = run_gpu(case, p, args.steps, args.warmup, args.timing_steps, args.scheme)
        log(f"  fp{p}: {res[p]['mlups_feq']:.1f} MLUPS feq, {res[p]['ms_per_step']:.4f} ms/step")
    head = 32 if 32 in res else precs[0]
    r = res[head]
    # dominant kernel: the fused collide+stream launch of the level with the
    # most cells (the finest level of this workload)
    dom_level = max(r["kernel_levels"], key=lambda lv: r["kernel_levels"][lv]["cells_per_launch"])
    dom = r["kernel_levels"][dom_level]
    kb = BYTES_PER_UPDATE[head] * dom["cells_per_launch"]
    achieved = kb / (dom["avg_launch_us"] * 1e-6) / 1e9
    synth = None
    if not args.no_synthetic and args.workload == "cylinder":
        log("synthetic 3-level AMR case")
        sc = build_case("synthetic")
        sres = {}
        for p in precs:
            log(f"gpu fp{p} synthetic: {args.synthetic_steps} coarse steps")
            sres[p] = run_gpu(sc, p, args.synthetic_steps, 5, min(args.timing_steps, 20), args.scheme)
            log(f"  fp{p}: {sres[p]['mlups_feq']:.1f} MLUPS feq, {sres[p]['ms_per_step']:.4f} ms/step")
        sh = sres[head]
        sl = max(sh["kernel_levels"], key=lambda lv: sh["kernel_levels"][lv]["cells_per_launch"])
        sd = sh["kernel_levels"][sl]
        sach = BYTES_PER_UPDATE[head] * sd["cells_per_launch"] / (sd["avg_launch_us"] * 1e-6) / 1e9
        lv = sc.topo.level
        synth = {
            "workload": ("unit cube, 8^3 level-0 bricks of 8^3 cells, 3 levels refined where a seeded "
                         "sum of 8 Gaussian blobs is largest (50% of the volume at level >= 1, 25% at "
                         "level 2, 2:1 balanced), divergence-free Fourier initial velocity |k| <= 4, "
                         "max |u| 0.01, bounce-back walls, tau 0.6"),
            "blocks_per_level": [int((lv == l).sum()) for l in range(int(lv.max()) + 1)],
            "autotune": {f"fp{p}": sres[p]["autotune"] for p in sres},
            "fine_equivalent_updates_per_step": sh["feq_per_step"],
            "actual_updates_per_step": sh["actual_per_step"],
            "fp32_mlups": round(sres[32]["mlups_feq"], 3) if 32 in sres else None,
            "fp64_mlups": round(sres[64]["mlups_feq"], 3) if 64 in sres else None,
            "ms_per_step": round(sh["ms_per_step"], 6),
            "roofline": {"bound": "hbm", "achieved": round(sach, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(sach / HBM_PEAK_GBS, 4),
                         "kernel": f"fused collide+stream (KBlock), level {sl}",
                         "cells_per_launch": sd["cells_per_launch"],
                         "avg_launch_us": round(sd["avg_launch_us"], 3)},
        }
    ens = None
    if not args.no_ensemble and args.workload == "cylinder":
        log("cavity ensemble")
        ens = run_ensemble(precs, args.ensemble_steps, 5, min(args.timing_steps, 10), scheme=args.scheme)
        for p in precs:
            log(f"  fp{p}: {ens[f'fp{p}']['mlups']:.1f} MLUPS, {ens[f'fp{p}']['ms_per_step']:.4f} ms/step")
    links = None
    if not args.no_links and args.workload == "cylinder":
        log("bouzidi link build")
        links = run_link_build(not args.no_cpu)
        log(f"  voxelized {links['voxelize']['cells']} cells in {links['voxelize']['gpu_ms']} ms "
            f"({links['voxelize']['exact_fallback_cells']} by the exact fallback); "
            f"{links['links']} links from {links['triangles']} triangles in {links['gpu_ms']} ms")
    mig = None
    if not args.no_migrate and args.workload == "cylinder":
        log("state migration after a forest change")
        mig = run_migrate(head)
        log(f"  {mig['blocks'][0]} -> {mig['blocks'][1]} blocks: {mig['forward_ms']} ms "
            f"({mig['forward_gbs']} GB/s), back {mig['back_ms']} ms ({mig['back_gbs']} GB/s)")
    adjo = adjc = None
    if not args.no_adjoint and args.workload == "cylinder":
        log("adjoint")
        adjo = run_adjoint(args.adjoint_steps, args.adjoint_checkpoint)
        log("adjoint of the 2-level channel")
        adjc = run_adjoint_channel(args.channel_adjoint_steps, args.adjoint_checkpoint)
        log(f"  {adjc['coarse_steps']} coarse steps: adjoint {adjc['adjoint_ms']} ms, ratio "
            f"{adjc['adjoint_over_forward_fp64']} (FP32 forward {adjc['fp32_forward']['over_forward_fp64']})")
        log(f"  FP32 forward / FP64 adjoint: {adjo['fp32_forward']['adjoint_ms']} ms -> "
            f"{adjo['fp32_forward']['over_forward_fp64']}x FP64 forward, errors {adjo['fp32_forward']['rel_err_vs_fp64_adjoint']}")
        log(f"  {adjo['steps']} steps: adjoint {adjo['adjoint_ms']} ms (forward sweep {adjo['forward_sweep_ms']}), "
            f"forward FP64 {adjo['forward_fp64_ms_per_update']} ms/update -> ratio {adjo['adjoint_over_forward_fp64']}")
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
            "workload": {
                "cylinder": ("cylinder channel, 2-level AMR, coarse 128x32x32 in 8^3 blocks, fine box "
                             "2D upstream to 6D downstream, D=16 fine cells, Bouzidi links, inlet "
                             "u=0.04, pressure outlet, tau=0.55"),
                "cavity": "lid-driven cavity 64^3, one padded block (halo 3), tau=0.6, lid 0.05",
                "synthetic": ("unit cube, 8^3 level-0 bricks, 3 levels refined on a Gaussian-blob sum, "
                              "divergence-free Fourier start, bounce-back walls, tau 0.6"),
            }[args.workload],
            "blocks": case.topo.num_blocks,
            "levels": int(case.topo.max_level) + 1,
            "actual_updates_per_step": r["actual_per_step"],
            "fine_equivalent_updates_per_step": r["feq_per_step"],
            "timed_region": "hipGraph replay of whole coarse steps",
            "autotune": r["autotune"],
        },
        "mlups_actual": round(r["mlups_actual"], 3),
        "fp32_mlups": round(res[32]["mlups_feq"], 3) if 32 in res else None,
        "fp64_mlups": round(res[64]["mlups_feq"], 3) if 64 in res else None,
        "fp32_over_fp64": (round(res[32]["mlups_feq"] / res[64]["mlups_feq"], 3)
                           if 32 in res and 64 in res else None),
        "adjoint_over_forward": adjoint_ratio(adjc),
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic_bytes(args.workload, head),
            "kernel": f"fused collide+stream (KBlock), level {dom_level}",
            "cells_per_launch": dom["cells_per_launch"],
            "avg_launch_us": round(dom["avg_launch_us"], 3),
            "bytes_per_update": BYTES_PER_UPDATE[head],
            "per_level_avg_launch_us": {str(k): round(v["avg_launch_us"], 3)
                                        for k, v in r["kernel_levels"].items()},
        },
        "cpu_baseline": cpu,
        "synthetic_amr": synth,
        "ensemble": ens,
        "link_build": links,
        "migration": mig,
        "adjoint": adjc,
        "adjoint_single_level": adjo,
        "drag_query_us": {f"fp{p}": (round(res[p]["drag_query_us"], 2) if res[p]["drag_query_us"] else None)
                          for p in res},
        "drag_fp32": res[32]["drag"] if 32 in res else None,
        "drag_fp64": res[64]["drag"] if 64 in res else None,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
