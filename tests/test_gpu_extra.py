"""FP32-shifted path accuracy against the FP64 oracle, the halo view of
block downloads, and the kernel-replay timing hook.

Accuracy bounds (FP32 populations, FP64 drag / mass sums): velocity L2
relative error <= 1e-5 and drag relative error <= 1e-4 against the FP64
oracle after the same number of coarse steps.
"""
import numpy as np
import pytest

from octlb_amd import abi, cases
from tests.util import make_gpu, make_oracle, random_pops

pytestmark = pytest.mark.gpu

VEL_L2_TOL = 1e-5
DRAG_TOL = 1e-4


def _two_level(ramp=0.0):
    return cases.cylinder_channel(bricks_x=2, base_level=1, fine=True, xc=0.5, radius=0.1,
                                  refine_lo=0.25, refine_hi=1.0, ramp_end_t=ramp)


def _velocity(core, fluid):
    rho = 1.0 + core.sum(-1)
    u = (core @ cases.C.astype(np.float64)) / rho[..., None]
    return u[fluid]


def test_fp32_accuracy_vs_fp64_oracle():
    case = _two_level()
    steps = 120
    o = make_oracle(case)
    o.set_threads(8)
    g = make_gpu(case, precision=32)
    o.advance(steps)
    g.advance(steps)
    fluid = case.kinds == abi.KIND_BULK
    u64 = _velocity(o.download_core(), fluid)
    u32 = _velocity(g.download_core(), fluid)
    l2 = np.linalg.norm(u32 - u64) / np.linalg.norm(u64)
    f64, f32 = o.drag(case.drag_bbox), g.drag(case.drag_bbox)
    drag_rel = np.linalg.norm(f32 - f64) / np.linalg.norm(f64)
    mass_rel = abs(g.mass() - o.mass()) / o.mass()
    print(f"fp32 vs fp64 after {steps} steps: vel L2 {l2:.3e}, drag {drag_rel:.3e}, mass {mass_rel:.3e}")
    assert l2 <= VEL_L2_TOL
    assert drag_rel <= DRAG_TOL
    assert mass_rel <= 1e-6


def _exchanged_slots(topo, edges, b):
    """(padded slot index, neighbour block, neighbour cell) of every halo
    slot of block b that a same-level face or edge exchange fills."""
    n = topo.n
    same = {(int(a), int(d)): int(r) for a, d, r in topo.same_level}
    diag = {}
    for a, d1, d2, dd in edges:
        if int(a) == b:
            diag[frozenset((int(d1), int(d2)))] = int(dd)
    out = []
    NP = n + 2
    for px in range(NP):
        for py in range(NP):
            for pz in range(NP):
                c = (px - 1, py - 1, pz - 1)
                outside = [ax for ax in range(3) if not 0 <= c[ax] < n]
                if not outside or len(outside) == 3:
                    continue
                dirs = [2 * ax + (1 if c[ax] >= n else 0) for ax in outside]
                nb = same.get((b, dirs[0])) if len(dirs) == 1 else diag.get(frozenset(dirs))
                if nb is None:
                    continue
                w = [(n - 1 if v < 0 else (0 if v >= n else v)) for v in c]
                out.append(((px * NP + py) * NP + pz, nb, w))
    return out


def test_download_block_halo_view():
    """Core and non-exchanged halo slots equal the oracle's bit for bit;
    exchanged halo slots hold the neighbour's boundary layer (what the next
    ghost exchange writes there)."""
    case = cases.cylinder_channel(bricks_x=4, bricks_yz=2, base_level=0, fine=False, xc=0.5,
                                  radius=0.15)
    pops = random_pops(case, 21)
    o = make_oracle(case, pops)
    g = make_gpu(case, 64, pops)
    o.advance(3)
    g.advance(3)
    core = o.download_core()
    edges = o.edges()
    checked = 0
    for b in (0, 5, case.topo.num_blocks - 1):
        gb = g.download_block(b).reshape(-1, 19)
        ob = o.download_block(b).reshape(-1, 19)
        ex = _exchanged_slots(case.topo, edges, b)
        exchanged = np.zeros(gb.shape[0], bool)
        for s, nb, w in ex:
            exchanged[s] = True
            np.testing.assert_array_equal(gb[s], core[nb, w[0], w[1], w[2]])
        np.testing.assert_array_equal(gb[~exchanged], ob[~exchanged])
        checked += len(ex)
    assert checked > 0


def test_time_kernel_leaves_state_unchanged():
    case = _two_level()
    g1 = make_gpu(case, 64)
    g2 = make_gpu(case, 64)
    g1.advance(3)
    for lv in sorted(set(int(x) for x in case.topo.level)):
        ms, cells = g1.time_kernel(lv, 3)
        assert ms > 0.0 and cells == np.sum(case.topo.level == lv) * case.n ** 3
    g1.advance(2)
    g2.advance(5)
    np.testing.assert_array_equal(g1.download_core(), g2.download_core())
    np.testing.assert_array_equal(g1.drag(case.drag_bbox), g2.drag(case.drag_bbox))


def test_snapshot_switch_matches_create_time_flag():
    # octlb_dev_set_snapshot (TimeLoop::set_snapshot_post_collide) on a
    # handle created without the snapshot gives the drag of a handle created
    # with it; switching it off and on again in mid-run as well
    import dataclasses
    case = _two_level()
    assert case.snapshot == 1
    a = make_gpu(case, 64)
    b = make_gpu(dataclasses.replace(case, snapshot=0), 64)
    b.set_snapshot(True)
    a.advance(3)
    b.advance(3)
    np.testing.assert_array_equal(b.drag(case.drag_bbox), a.drag(case.drag_bbox))
    b.set_snapshot(False)
    a.advance_graph(2)
    b.advance_graph(2)
    b.set_snapshot(True)
    a.advance_graph(17)
    b.advance_graph(17)
    np.testing.assert_array_equal(b.download_core(), a.download_core())
    np.testing.assert_array_equal(b.drag(case.drag_bbox), a.drag(case.drag_bbox))
    a.close()
    b.close()
