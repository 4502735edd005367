"""The two update schemes of the fused kernel give bit-identical states.

Variant 0: every 8^3 block pulls through its LDS box (cells + one-cell
apron, the apron re-collided).  Variant 1: blocks whose cells all stream
plainly and whose 18 halo parts are exchanged at their level push their
post-collide values (push tiles of two blocks), the rest pull and push their
surface values into push-class neighbours.  Same arithmetic, only the data
movement differs, so populations, drag and mass agree bit for bit in FP64 and
FP32; the autotuner leaves the state unchanged.
"""
import numpy as np
import pytest

from octlb_amd import abi, cases
from octlb_amd.solver import GpuSolver

pytestmark = pytest.mark.gpu


def _cases():
    return {
        "cylinder-small": cases.cylinder_channel(bricks_x=2, base_level=1, fine=True, xc=0.5, radius=0.1,
                                                 refine_lo=0.25, refine_hi=1.0),
        "cylinder-headline": cases.cylinder_channel(bricks_x=16, bricks_yz=4, base_level=0, fine=True, xc=1.0,
                                                    radius=0.125, u_in=0.04, tau=0.55),
        "synthetic-small": cases.synthetic_amr(bricks=4, max_level=2, frac=(0.5, 0.25), seed=3),
    }


def _run(case, precision, variant, steps):
    sim = case.apply(GpuSolver(case.topo, case.params(precision)))
    sim.set_kernel_variant(variant)
    v, npush = sim.kernel_variant()
    sim.advance(3)
    sim.advance_graph(steps)
    out = (sim.download_core(), sim.drag(case.drag_bbox) if case.drag_bbox is not None else np.zeros(3),
           sim.mass(), npush)
    sim.close()
    return out


@pytest.mark.parametrize("name", ["cylinder-small", "cylinder-headline", "synthetic-small"])
@pytest.mark.parametrize("precision", [64, 32])
def test_push_and_pull_variants_agree(name, precision):
    case = _cases()[name]
    a = _run(case, precision, 0, 21)
    b = _run(case, precision, 1, 21)
    print(f"{name} fp{precision}: {b[3]} of {case.topo.num_blocks} blocks push", flush=True)
    assert a[3] == 0
    assert b[3] > 0 or name == "cylinder-small"  # its blocks all touch a wall, a face of another level or the body
    np.testing.assert_array_equal(b[0], a[0])
    np.testing.assert_array_equal(b[1], a[1])
    assert b[2] == a[2]


def test_push_variant_ensemble_with_fast_stats():
    ens = cases.cavity_ensemble(count=6, sizes=(32, 48), seed=4)
    outs = []
    for v in (0, 1):
        s = GpuSolver.ensemble(ens, 32)
        s.set_bc_kinds(ens.kinds)
        s.init_equilibrium(1.0, (0.0, 0.0, 0.0), 0xFF)
        s.set_kernel_variant(v)
        s.advance_graph(20)
        outs.append((s.download_core(), [s.member_average_rho(m) for m in range(6)], s.kernel_variant()[1]))
        s.close()
    assert outs[0][2] == 0 and outs[1][2] > 0
    np.testing.assert_array_equal(outs[1][0], outs[0][0])
    assert outs[1][1] == outs[0][1]


def test_autotune_keeps_state_and_picks_a_variant():
    case = _cases()["cylinder-headline"]
    sim = case.apply(GpuSolver(case.topo, case.params(32)))
    sim.advance(2)
    before = sim.download_core()
    chosen, ms = sim.autotune(reps=5)
    print(f"autotune: variant {chosen}, fused-kernel ms per coarse step {ms}", flush=True)
    assert chosen in (0, 1) and ms[0] > 0 and ms[1] > 0
    np.testing.assert_array_equal(sim.download_core(), before)
    ref = case.apply(GpuSolver(case.topo, case.params(32)))
    ref.set_kernel_variant(1 - chosen)
    ref.advance(2)
    sim.advance_graph(5)
    ref.advance_graph(5)
    np.testing.assert_array_equal(sim.download_core(), ref.download_core())
    sim.close()
    ref.close()


def test_coupling_epilogues_are_bit_identical():
    """The half-time prolongation and the restriction run inside the update
    launches or as separate kernels: the 2-level channel's FP64 and FP32
    states after 12 coarse steps (eager) and after 16 more (graph replay) are
    the same bit for bit, and the autotuner leaves the state unchanged."""
    case = cases.cylinder_channel(bricks_x=2, base_level=1, fine=True, xc=0.5, radius=0.1,
                                  refine_lo=0.25, refine_hi=1.0)
    for prec in (64, 32):
        out = {}
        for mode in (0, 3):
            sim = case.apply(GpuSolver(case.topo, case.params(prec)))
            sim.set_coupling_epilogues(mode)
            assert sim.coupling_epilogues()[0] == mode
            sim.advance(12)
            a = sim.download_core()
            sim.prepare_graphs()
            sim.advance_graph(16)
            out[mode] = (a, sim.download_core(), sim.drag(case.drag_bbox))
            if mode == 3:
                before = sim.download_core()
                sim.autotune(reps=2)
                np.testing.assert_array_equal(sim.download_core(), before)
                print(f"fp{prec}: autotuned epilogues {sim.coupling_epilogues()}", flush=True)
            sim.close()
        for k in range(3):
            np.testing.assert_array_equal(out[0][k], out[3][k])
