"""Multi-level discrete adjoint (two refinement levels with the Lagrava
coupler, FD velocity inlet and pressure outlet, Bouzidi cylinder) on a
reduced 2-level channel: 72 blocks, 24 coarse steps, J = mean x-drag over
coarse steps 13..24, checkpoint every 5 steps, inlet ramp over 8 steps.

* Its forward sweep is the production coarse step bit for bit (FP64 core
  populations after octlb_dev_advance) and J equals the production drag.
* dJ/du_in (through the velocity-Dirichlet collide and the FD pass, whose
  stencils read the prescribed velocity) and dJ/dq (through the Bouzidi
  rows of the fine level's stream, whose adjoint crosses the coupler and the
  FD pass) match central differences of the FP64 oracle's forward
  (eps = 1e-6) to 1e-4: u_in's x component, the three largest |dJ/dq|, and
  a random direction over all links.
* With FP32 forward states (FP64 adjoint) the same gradients match the same
  central differences to 1e-4.
* The gradients do not depend on the checkpoint interval, nor on whether a
  segment's tapes are kept for all its steps (the default here) or each step
  is re-run just before its reverse (a zero tape budget).
"""
import numpy as np
import pytest

from octlb_amd import abi, cases
from octlb_amd.adjoint import DragAdjoint
from octlb_amd.solver import GpuSolver
from oracle.binding import OracleSim

pytestmark = pytest.mark.gpu

STEPS, WIN, CK = 24, (12, 24), 5
EPS = 1e-6
TOL = 1e-4


@pytest.fixture(scope="module")
def channel():
    return cases.cylinder_channel(bricks_x=2, base_level=1, fine=True, xc=0.5, radius=0.1, refine_lo=0.25,
                                  refine_hi=1.0, u_in=0.04, ramp_end_t=8.0)


def _u_in(case):
    for s in case.specs:
        if s.inlet != abi.INLET_NONE:
            return np.array([s.inlet_u[0], s.inlet_u[1], s.inlet_u[2]])
    raise AssertionError("no inlet")


_FD_CACHE = {}


def _central_difference(case, key, plus, minus):
    """(J(plus) - J(minus)) / 2 eps of the FP64 oracle, computed once per key."""
    if key not in _FD_CACHE:
        _FD_CACHE[key] = (_oracle_J(case, **plus) - _oracle_J(case, **minus)) / (2 * EPS)
    return _FD_CACHE[key]


def _oracle_J(case, q=None, u_in=None):
    adj = DragAdjoint(case, STEPS, WIN)
    p = adj._params(u_in)
    s = OracleSim(case.topo, p)
    s.set_bc_kinds(case.kinds)
    s.set_links(case.link, case.q if q is None else q)
    s.init_equilibrium(1.0, case.init_u0, case.init_mask)
    s.set_threads(8)
    J = 0.0
    for st in range(1, STEPS + 1):
        s.advance(1)
        if WIN[0] < st <= WIN[1]:
            J += s.drag(case.drag_bbox)[0]
    s.close()
    return J / (WIN[1] - WIN[0])


@pytest.fixture(scope="module")
def device(channel):
    print("multi-level adjoint: 72 blocks, 24 coarse steps, checkpoint 5", flush=True)
    adj = DragAdjoint(channel, STEPS, WIN, checkpoint=CK)
    J, du0, dq, core = adj.evaluate(final_core=True)
    assert adj.packed, "the default tape budget should hold this case's segments"
    return J, dq, adj.inlet_grad.copy(), core


def test_forward_is_the_production_step(channel, device):
    J, _, _, core = device
    sol = GpuSolver(channel.topo, channel.params(64))
    sol.set_bc_kinds(channel.kinds)
    sol.set_links(channel.link, channel.q)
    sol.init_equilibrium(1.0, channel.init_u0, channel.init_mask)
    Jp = 0.0
    for st in range(1, STEPS + 1):
        sol.advance(1)
        if WIN[0] < st <= WIN[1]:
            Jp += sol.drag(channel.drag_bbox)[0]
    Jp /= WIN[1] - WIN[0]
    ref = sol.download_core()
    sol.close()
    print(f"J {J:.15e} vs production {Jp:.15e}", flush=True)
    np.testing.assert_array_equal(core, ref)
    assert J == pytest.approx(Jp, rel=1e-12)


def test_inlet_velocity_gradient_matches_central_differences(channel, device):
    _, _, gi, _ = device
    u = _u_in(channel)
    e = np.array([EPS, 0.0, 0.0])
    fd = _central_difference(channel, "u_in_x", dict(u_in=u + e), dict(u_in=u - e))
    print(f"dJ/du_in,x adjoint {gi[0]:.10e} vs central difference {fd:.10e}", flush=True)
    assert abs(fd) > 0
    assert abs(gi[0] - fd) <= TOL * abs(fd)


def _direction(q):
    rng = np.random.default_rng(5)
    v = rng.normal(size=q.shape)
    return v / np.abs(v).max()


def test_wall_fraction_gradients_match_central_differences(channel, device):
    _, dq, _, _ = device
    q = np.array(channel.q, dtype=np.float64)
    top = np.argsort(-np.abs(dq))[:3]
    for k in top:
        qp, qm = q.copy(), q.copy()
        qp[k] += EPS
        qm[k] -= EPS
        fd = _central_difference(channel, ("q", int(k)), dict(q=qp), dict(q=qm))
        print(f"link {k}: dJ/dq adjoint {dq[k]:.10e} vs central difference {fd:.10e}", flush=True)
        assert abs(dq[k] - fd) <= TOL * abs(fd)
    v = _direction(q)
    fd = _central_difference(channel, "q_dir", dict(q=q + EPS * v), dict(q=q - EPS * v))
    ad = float(dq @ v)
    print(f"random direction: adjoint {ad:.10e} vs central difference {fd:.10e}", flush=True)
    assert abs(ad - fd) <= TOL * abs(fd)


def test_checkpoint_interval_does_not_change_the_gradients(channel, device):
    J, dq, gi, _ = device
    adj = DragAdjoint(channel, STEPS, WIN, checkpoint=STEPS)
    J2, _, dq2 = adj.evaluate()
    assert J2 == J
    np.testing.assert_array_equal(dq2, dq)
    np.testing.assert_array_equal(adj.inlet_grad, gi)


def test_recompute_mode_matches_the_packed_tapes(channel, device):
    J, dq, gi, _ = device
    adj = DragAdjoint(channel, STEPS, WIN, checkpoint=CK, tape_budget=0)
    J2, _, dq2 = adj.evaluate()
    assert not adj.packed
    assert J2 == J
    np.testing.assert_array_equal(dq2, dq)
    np.testing.assert_array_equal(adj.inlet_grad, gi)


def test_fp32_forward_gradients_match_central_differences(channel, device):
    """FP32 states in the forward sweep and the taping pass (the FP32 path's
    arithmetic), adjoint in FP64: dJ/du_in, the three largest |dJ/dq| and
    the random direction against the FP64 oracle's central differences."""
    _, dq64, _, _ = device
    adj = DragAdjoint(channel, STEPS, WIN, checkpoint=CK, precision=32)
    _, _, dq = adj.evaluate()
    gi = adj.inlet_grad
    u = _u_in(channel)
    e = np.array([EPS, 0.0, 0.0])
    fd = _central_difference(channel, "u_in_x", dict(u_in=u + e), dict(u_in=u - e))
    print(f"FP32 forward: dJ/du_in,x {gi[0]:.10e} vs central difference {fd:.10e}", flush=True)
    assert abs(gi[0] - fd) <= TOL * abs(fd)
    q = np.array(channel.q, dtype=np.float64)
    for k in np.argsort(-np.abs(dq64))[:3]:
        qp, qm = q.copy(), q.copy()
        qp[k] += EPS
        qm[k] -= EPS
        fd = _central_difference(channel, ("q", int(k)), dict(q=qp), dict(q=qm))
        print(f"FP32 forward: link {k} dJ/dq {dq[k]:.10e} vs central difference {fd:.10e}", flush=True)
        assert abs(dq[k] - fd) <= TOL * abs(fd)
    v = _direction(q)
    fd = _central_difference(channel, "q_dir", dict(q=q + EPS * v), dict(q=q - EPS * v))
    ad = float(dq @ v)
    print(f"FP32 forward: random direction {ad:.10e} vs central difference {fd:.10e}", flush=True)
    assert abs(ad - fd) <= TOL * abs(fd)
