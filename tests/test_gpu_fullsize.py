"""Full-size checks of the benchmarked configurations on the GPU.

* The 2-level cylinder channel exactly as the benchmark builds it (coarse
  128x32x32 in 8^3 blocks, one refinement level, D = 16 fine cells, Bouzidi
  links, velocity inlet, pressure outlet, tau 0.55; 1,152 blocks): FP64 GPU
  against the FP64 oracle (OpenMP over blocks) after 40 coarse steps --
  populations, drag and mass to 1e-12 relative -- and FP32 against FP64
  after 2,000 coarse steps, when the flow has passed the cylinder and the
  wake has formed: velocity L2 <= 1e-5 and drag <= 1e-4 relative.
* The synthetic 3-level AMR case (9,689 blocks) and the 64-cavity ensemble
  (49,848 blocks, 25.5 M cells) at full size: FP32 against FP64 after 50
  coarse steps (velocity L2 <= 1e-5, per member for the ensemble) and the
  mass drift of both precisions.
"""
import os

import numpy as np
import pytest

from octlb_amd import abi, cases
from octlb_amd.solver import GpuSolver

pytestmark = pytest.mark.gpu

VEL_L2_TOL = 1e-5
DRAG_TOL = 1e-4


def _headline():
    # the benchmark's workload (bench.py build_case("cylinder"))
    return cases.cylinder_channel(bricks_x=16, bricks_yz=4, base_level=0, fine=True, xc=1.0,
                                  radius=0.125, u_in=0.04, tau=0.55)


def _gpu(case, precision):
    sim = case.apply(GpuSolver(case.topo, case.params(precision)))
    sim.prepare_graphs()
    return sim


def _threads():
    env = os.environ.get("OMP_NUM_THREADS", "")
    return int(env) if env.isdigit() and int(env) > 0 else min(16, len(os.sched_getaffinity(0)))


def _vel_l2(u32, u64, fluid):
    return float(np.linalg.norm(u32[fluid] - u64[fluid]) / np.linalg.norm(u64[fluid]))


def test_headline_fp64_matches_oracle():
    from oracle.binding import OracleSim
    case = _headline()
    assert case.topo.num_blocks == 1152
    steps = 40
    g = _gpu(case, 64)
    o = case.apply(OracleSim(case.topo, case.params(64)))
    o.set_threads(_threads())
    g.advance_graph(steps)
    o.advance(steps)
    pg, po = g.download_core(), o.download_core()
    err = float(np.abs(pg - po).max() / np.abs(po).max())
    dg, do = g.drag(case.drag_bbox), o.drag(case.drag_bbox)
    derr = float(np.abs(dg - do).max() / np.abs(do).max())
    mg, mo = g.mass(), o.mass()
    print(f"headline FP64 vs oracle after {steps} steps: pops {err:.3e}, drag {derr:.3e} "
          f"({do}), mass {abs(mg - mo) / mo:.3e}", flush=True)
    assert np.abs(do).max() > 0.0
    assert err <= 1e-12
    assert derr <= 1e-12
    assert abs(mg - mo) <= 1e-12 * mo
    g.close()
    o.close()


def test_headline_fp32_vs_fp64_developed_flow():
    case = _headline()
    fluid = case.kinds == abi.KIND_BULK
    sims = {p: _gpu(case, p) for p in (32, 64)}
    done = 0
    for target in (100, 500, 2000):
        for s in sims.values():
            s.advance_graph(target - done)
        done = target
        u = {}
        for p, s in sims.items():
            core = s.download_core()
            rho = 1.0 + core.sum(-1)
            u[p] = (core @ cases.C.astype(np.float64)) / rho[..., None]
        l2 = _vel_l2(u[32], u[64], fluid)
        d32, d64 = sims[32].drag(case.drag_bbox), sims[64].drag(case.drag_bbox)
        drel = float(np.linalg.norm(d32 - d64) / np.linalg.norm(d64))
        m32, m64 = sims[32].mass(), sims[64].mass()
        print(f"headline FP32 vs FP64 after {target} steps: vel L2 {l2:.3e}, drag {drel:.3e} "
              f"(fp64 {d64}), mass {abs(m32 - m64) / m64:.3e}", flush=True)
    # the wake has formed: the drag is the converged flow's, not the first
    # pressure wave's
    assert abs(d64[0]) > 1e-2
    assert l2 <= VEL_L2_TOL
    assert drel <= DRAG_TOL
    for s in sims.values():
        s.close()


def test_synthetic_three_level_full_size():
    case = cases.synthetic_amr(bricks=8, max_level=2, frac=(0.5, 0.25), seed=7)
    assert case.topo.num_blocks == 9689
    fluid = case.kinds == abi.KIND_BULK
    lv = case.topo.level
    w = (1.0 / 8.0) ** lv  # a cell's volume in level-0 cells
    out = {}
    for p in (32, 64):
        s = _gpu(case, p)
        core0 = s.download_core()
        m0 = float(((1.0 + core0.sum(-1)) * fluid).sum(axis=(1, 2, 3)) @ w)
        s.advance_graph(50)
        core = s.download_core()
        rho = 1.0 + core.sum(-1)
        u = (core @ cases.C.astype(np.float64)) / rho[..., None]
        m1 = float((rho * fluid).sum(axis=(1, 2, 3)) @ w)
        out[p] = (u, m0, m1)
        s.close()
    l2 = _vel_l2(out[32][0], out[64][0], fluid)
    drift = {p: abs(out[p][2] - out[p][1]) / out[p][1] for p in (32, 64)}
    print(f"synthetic 3-level after 50 steps: vel L2 {l2:.3e}, volume-weighted mass drift "
          f"FP64 {drift[64]:.3e} FP32 {drift[32]:.3e}", flush=True)
    assert l2 <= VEL_L2_TOL
    # closed walls: mass changes only through the coupler's interpolation
    assert drift[64] <= 1e-4
    assert abs(drift[32] - drift[64]) <= 1e-6


def test_ensemble_full_size():
    ens = cases.cavity_ensemble(count=64)
    assert ens.topo.num_blocks == 49848
    res = {}
    for p in (32, 64):
        s = GpuSolver.ensemble(ens, p)
        s.set_bc_kinds(ens.kinds)
        s.init_equilibrium(1.0, (0.0, 0.0, 0.0), 0xFF)
        s.prepare_graphs()
        s.advance_graph(50)
        rho, u = s.macro_fields()
        avg = [s.member_average_rho(m) for m in range(len(ens.members))]
        res[p] = (rho, u, avg)
        s.close()
    worst, worst_m = 0.0, -1
    for m in range(len(ens.members)):
        sl = slice(int(ens.member_begin[m]), int(ens.member_begin[m + 1]))
        u32, u64 = res[32][1][sl], res[64][1][sl]
        l2 = float(np.linalg.norm(u32 - u64) / np.linalg.norm(u64))
        if l2 > worst:
            worst, worst_m = l2, m
    drift = max(abs(a - 1.0) for a in res[64][2])
    adiff = max(abs(a - b) for a, b in zip(res[32][2], res[64][2]))
    print(f"ensemble after 50 steps: worst member vel L2 {worst:.3e} (member {worst_m}), "
          f"ConstRho average |rho-1| FP64 {drift:.3e}, FP32-FP64 {adiff:.3e}", flush=True)
    assert worst <= VEL_L2_TOL
    assert adiff <= 1e-6
