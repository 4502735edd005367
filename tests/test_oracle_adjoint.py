"""Pins the oracle's discrete adjoint (ol_adjoint) of the single-level step
by central differences of the oracle's own forward (eps = 1e-6), the check
the adjoint has to pass: the reference has no adjoint, so its gradients are
defined only through the forward it differentiates.

Case: a closed 16-block box (4 x 2 x 2 bricks of 8^3 cells, bounce-back
walls on every face, so faces and edges of the halo exchange are crossed),
a z-spanning cylinder with exact ray-circle Bouzidi fractions on 752 links
(both branches of the Bouzidi formula occur), fluid started at
feq(1, (0.05, 0, 0)), tau 0.6; J = mean x-drag over steps 11..30.

A plausible slip -- the Bouzidi derivative of the other branch, the
interior population taken from the wrong slot, the exchange transposed
without clearing the overwritten halo, the u-derivative of feq without its
-3 u_d term, rho instead of 1/rho in du/df, the drag's live term on c_i
instead of c_opp(i) -- moves at least one of these gradients far outside
the tolerance.
"""
import numpy as np
import pytest

from octlb_amd import cases
from oracle.binding import OracleSim

STEPS, WIN = 30, (10, 30)
EPS = 1e-6
DIR = np.array([1.0, 0.0, 0.0])


@pytest.fixture(scope="module")
def box():
    return cases.cylinder_channel(bricks_x=4, base_level=0, fine=False, bricks_yz=2, xc=0.6, radius=0.125,
                                  u_in=0.05, tau=0.6, open_ends=False)


def _sim(case, q=None):
    s = OracleSim(case.topo, case.params(64))
    s.set_bc_kinds(case.kinds)
    s.set_links(case.link, case.q if q is None else q)
    return s


def _J(case, u0=None, q=None):
    u0 = case.init_u0 if u0 is None else u0
    return _sim(case, q).adjoint(STEPS, WIN, DIR, case.drag_bbox, u0, link=case.link[:0])[0]


@pytest.fixture(scope="module")
def grad(box):
    return _sim(box).adjoint(STEPS, WIN, DIR, box.drag_bbox, box.init_u0, link=box.link)


def test_objective_is_the_mean_drag(box, grad):
    s = _sim(box)
    s.init_equilibrium(1.0, box.init_u0, box.init_mask)
    acc = 0.0
    for st in range(1, STEPS + 1):
        s.advance(1)
        if WIN[0] < st <= WIN[1]:
            acc += s.drag(box.drag_bbox)[0]
    assert grad[0] == pytest.approx(acc / (WIN[1] - WIN[0]), rel=1e-14, abs=0)


def test_inlet_velocity_gradient_matches_central_differences(box, grad):
    _, du, _ = grad
    for d in range(3):
        up = np.array(box.init_u0, float)
        um = up.copy()
        up[d] += EPS
        um[d] -= EPS
        fd = (_J(box, u0=up) - _J(box, u0=um)) / (2 * EPS)
        assert du[d] == pytest.approx(fd, rel=1e-6, abs=1e-7), d


def test_link_fraction_gradients_match_central_differences(box, grad):
    _, _, dq = grad
    q = box.q
    picks = []
    for branch in (q >= 0.5 + 1e-4, q < 0.5 - 1e-4):
        idx = np.nonzero(branch)[0]
        assert idx.size > 0, "both Bouzidi branches must occur"
        picks += list(idx[np.argsort(-np.abs(dq[idx]))[:2]])
    for k in picks:
        qp, qm = q.copy(), q.copy()
        qp[k] += EPS
        qm[k] -= EPS
        fd = (_J(box, q=qp) - _J(box, q=qm)) / (2 * EPS)
        assert dq[k] == pytest.approx(fd, rel=1e-4, abs=1e-8), (k, q[k])


def test_directional_derivative_over_all_links(box, grad):
    _, _, dq = grad
    rng = np.random.default_rng(5)
    v = rng.standard_normal(box.q.size)
    # keep every link inside its branch
    room = np.minimum(np.abs(box.q - 0.5), np.minimum(box.q, 1.0 - box.q))
    v *= np.minimum(1.0, room / (10 * EPS))
    fd = (_J(box, q=box.q + EPS * v) - _J(box, q=box.q - EPS * v)) / (2 * EPS)
    assert float(dq @ v) == pytest.approx(fd, rel=1e-5)
