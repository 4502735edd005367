"""GPU (HIP, gfx950) path vs the CPU FP64 oracle on identical inputs.

The FP64 GPU build uses no FMA contraction and the reference's evaluation
order, so most cases are compared bit for bit; long runs and AMR runs use the
1e-12 relative bound of the north-star contract.  The oracle itself is pinned
to the reference's own code in tests/test_oracle_vs_reference.py.
"""
import numpy as np
import pytest

from octlb_amd import abi, cases
from octlb_amd.forest import single_block_topology
from tests.util import make_gpu, make_oracle, random_pops, rel_err

pytestmark = pytest.mark.gpu


def _single_block_case():
    n, h = 8, 1
    topo = single_block_topology(n, h)
    kinds = np.zeros((1, n, n, n), np.uint8)
    specs = [abi.make_spec(abi.XMIN, abi.BC_INTERP_VELOCITY, inlet=abi.INLET_UNIFORM,
                           inlet_u=(0.03, 0.0, 0.0), ramp_end_t=4.0),
             abi.make_spec(abi.XMAX, abi.BC_INTERP_PRESSURE, rho_target=1.0)]
    cases.stamp_tree_boundary(kinds, topo.tree_boundary, specs)
    kinds[0, 3:5, 3:5, 3:5] = abi.KIND_SOLID
    kinds[0, 3:5, 3:5, 2] = abi.KIND_BOUZIDI
    kinds[0, 6, 1, 1] = abi.KIND_BOUNCE_BACK
    link = np.array([[0, (3 * n + 3) * n + 1, 12], [0, (4 * n + 4) * n + 1, 12],
                     [0, (3 * n + 4) * n + 1, 12]], np.int32)
    q = np.array([0.35, 0.7, 0.5])
    return cases.Case(name="block", topo=topo, omega=1 / 0.6, handler=abi.HANDLER_CONCRETE,
                      const_rho=0, bouzidi=1, snapshot=1, specs=specs, kinds=kinds, link=link, q=q,
                      drag_bbox=np.array([0.2, 0.2, 0.0, 0.8, 0.8, 1.0]))


def _run_both(case, steps, pops=None, precision=64):
    o = make_oracle(case, pops)
    g = make_gpu(case, precision, pops)
    o.advance(steps)
    g.advance(steps)
    return o, g


def test_single_block_bitwise():
    case = _single_block_case()
    pops = random_pops(case, 11)
    o, g = _run_both(case, 5, pops)
    np.testing.assert_array_equal(g.download_block(0), o.download_block(0))
    bb = case.drag_bbox
    np.testing.assert_array_equal(g.drag(bb), o.drag(bb))
    assert abs(g.mass() - o.mass()) <= 1e-12 * abs(o.mass())


def test_noop_handler_bitwise():
    case = _single_block_case()
    case.handler = abi.HANDLER_NOOP
    case.bouzidi = 0
    case.link = np.zeros((0, 3), np.int32)
    case.q = np.zeros(0)
    pops = random_pops(case, 12)
    o, g = _run_both(case, 4, pops)
    np.testing.assert_array_equal(g.download_block(0), o.download_block(0))


def test_cavity_padded_bitwise():
    case = cases.cavity(16, tau=0.6, u_lid=0.05)
    pops = random_pops(case, 13, scale=0.002)
    o, g = _run_both(case, 6, pops)
    np.testing.assert_array_equal(g.download_block(0), o.download_block(0))
    assert g.average_rho() == o.average_rho()


def test_cavity64_from_rest():
    case = cases.cavity(64, tau=0.6, u_lid=0.05)
    o, g = _run_both(case, 30)
    a, b = g.download_core(), o.download_core()
    assert rel_err(a, b) <= 1e-12
    assert abs(g.average_rho() - o.average_rho()) <= 1e-13


def test_cavity64_1000_steps():
    """The plain reference configuration itself: the 64^3 lid-driven cavity
    (tau 0.6, lid 0.05, halfway bounce-back walls) from rest for 1,000 steps,
    FP64, against the oracle (OpenMP on the host): populations and the
    ConstRho-free average density to 1e-12."""
    import os
    case = cases.cavity(64, tau=0.6, u_lid=0.05)
    o = make_oracle(case)
    o.set_threads(max(1, min(16, os.cpu_count() or 1)))
    g = make_gpu(case)
    print("cavity 64^3: 1,000 steps on the oracle and the device", flush=True)
    o.advance(1000)
    g.advance(1000)
    a, b = g.download_core(), o.download_core()
    err = rel_err(a, b)
    print(f"max relative population difference {err:.3e}", flush=True)
    assert err <= 1e-12
    assert abs(g.average_rho() - o.average_rho()) <= 1e-13
    assert abs(g.mass() - o.mass()) <= 1e-12 * o.mass()


def test_uniform_multiblock_bitwise():
    case = cases.cylinder_channel(bricks_x=4, bricks_yz=2, base_level=0, fine=False, xc=0.5,
                                  radius=0.15)
    pops = random_pops(case, 14)
    o, g = _run_both(case, 6, pops)
    np.testing.assert_array_equal(g.download_core(), o.download_core())
    np.testing.assert_array_equal(g.drag(case.drag_bbox), o.drag(case.drag_bbox))


def _two_level():
    return cases.cylinder_channel(bricks_x=2, base_level=1, fine=True, xc=0.5, radius=0.1,
                                  refine_lo=0.25, refine_hi=1.0)


def test_two_level_cylinder():
    case = _two_level()
    assert case.topo.max_level == 2 and case.topo.num_cf > 0
    o, g = _run_both(case, 12)
    a, b = g.download_core(), o.download_core()
    assert rel_err(a, b) <= 1e-12
    fd, fo = g.drag(case.drag_bbox), o.drag(case.drag_bbox)
    assert np.abs(fd - fo).max() <= 1e-12 * max(np.abs(fo).max(), 1e-30)
    assert abs(g.mass() - o.mass()) <= 1e-12 * o.mass()


def test_graph_replay_matches_eager():
    case = _two_level()
    g1 = make_gpu(case)
    g2 = make_gpu(case)
    g1.advance(5)
    g2.advance_graph(5)
    np.testing.assert_array_equal(g1.download_core(), g2.download_core())
    g1.advance(2)
    g2.advance_graph(2)
    np.testing.assert_array_equal(g1.download_core(), g2.download_core())
    # 16-step graphs (two replays) followed by single-step graphs
    g1.advance(35)
    g2.advance_graph(35)
    np.testing.assert_array_equal(g1.download_core(), g2.download_core())
    np.testing.assert_array_equal(g1.drag(case.drag_bbox), g2.drag(case.drag_bbox))


def test_fd_blocks_on_coarse_fine_interface():
    """Fine region covering part of the inlet: FD blocks are also coupling
    blocks, so the FD pass and the coupling steps run as separate launches."""
    case = cases.cylinder_channel(bricks_x=2, base_level=1, fine=True, xc=0.5, radius=0.1,
                                  refine_lo=0.0, refine_hi=0.75, refine_y=(0.0, 0.5))
    assert case.topo.num_cf > 0
    o, g = _run_both(case, 8)
    a, b = g.download_core(), o.download_core()
    assert rel_err(a, b) <= 1e-12
    fd, fo = g.drag(case.drag_bbox), o.drag(case.drag_bbox)
    assert np.abs(fd - fo).max() <= 1e-12 * max(np.abs(fo).max(), 1e-30)


def test_synthetic_three_level():
    """Synthetic 3-level octree (Gaussian-blob refinement, Fourier initial
    velocity, bounce-back walls), FP64 against the oracle."""
    case = cases.synthetic_amr(bricks=2, max_level=2, frac=(0.5, 0.25), seed=7)
    assert case.topo.max_level == 2
    o, g = _run_both(case, 4)
    a, b = g.download_core(), o.download_core()
    assert rel_err(a, b) <= 1e-12
    assert abs(g.mass() - o.mass()) <= 1e-12 * o.mass()


def test_two_level_cylinder_fine_region_interior():
    """Fine level away from inlet and outlet (no FD blocks of its own): the
    fine update launch before the full-time prolongation performs it."""
    case = cases.cylinder_channel(bricks_x=2, base_level=1, fine=True, xc=1.0, radius=0.1,
                                  refine_lo=0.6, refine_hi=1.4)
    lv = case.topo.level
    fd = [b for b in range(case.topo.num_blocks)
          if ((case.kinds[b] == abi.KIND_VELOCITY_DIRICHLET) | (case.kinds[b] == abi.KIND_PRESSURE_DIRICHLET)).any()]
    assert case.topo.num_cf > 0 and all(lv[b] == lv.min() for b in fd)
    o, g = _run_both(case, 12)
    assert rel_err(g.download_core(), o.download_core()) <= 1e-12
    fd_, fo = g.drag(case.drag_bbox), o.drag(case.drag_bbox)
    assert np.abs(fd_ - fo).max() <= 1e-12 * max(np.abs(fo).max(), 1e-30)
    g2 = make_gpu(case)
    g2.advance_graph(12)
    np.testing.assert_array_equal(g2.download_core(), g.download_core())
