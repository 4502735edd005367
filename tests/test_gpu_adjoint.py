"""Device discrete adjoint (octlb_adj_*) against the oracle's FP64 adjoint
(itself pinned by central differences of the oracle's forward) on the closed
16-block cylinder box: J, dJ/du0 and every dJ/dq to 1e-9 relative (the sums
run in another order); the adjoint's own forward sweep against
octlb_dev_advance bit for bit; gradients independent of the checkpoint
interval; the torch.autograd wrapper; central differences of the device
objective; the scope check of octlb_adj_create; and the FP32-forward /
FP64-adjoint mode: its forward sweep equals the FP32 production step bit for
bit and its gradients stay within 1e-4 of the FP64 ones.
"""
import ctypes as C

import numpy as np
import pytest
import torch

from octlb_amd import abi, cases
from octlb_amd.adjoint import DragAdjoint, DragObjective
from octlb_amd.solver import GpuSolver
from oracle.binding import OracleSim

pytestmark = pytest.mark.gpu

STEPS, WIN = 30, (10, 30)


@pytest.fixture(scope="module")
def box():
    return cases.cylinder_channel(bricks_x=4, base_level=0, fine=False, bricks_yz=2, xc=0.6, radius=0.125,
                                  u_in=0.05, tau=0.6, open_ends=False)


@pytest.fixture(scope="module")
def oracle_grad(box):
    s = OracleSim(box.topo, box.params(64))
    s.set_bc_kinds(box.kinds)
    s.set_links(box.link, box.q)
    return s.adjoint(STEPS, WIN, (1.0, 0.0, 0.0), box.drag_bbox, box.init_u0, link=box.link)


@pytest.fixture(scope="module")
def device_grad(box):
    print("device adjoint: 16-block box, 30 steps, checkpoint 7", flush=True)
    return DragAdjoint(box, STEPS, WIN, checkpoint=7).evaluate(final_core=True)


def test_matches_oracle_adjoint(oracle_grad, device_grad):
    Jo, duo, dqo = oracle_grad
    J, du, dq, _ = device_grad
    print(f"J {J:.15e} vs {Jo:.15e}; du {du} vs {duo}", flush=True)
    assert J == pytest.approx(Jo, rel=1e-12)
    np.testing.assert_allclose(du, duo, rtol=1e-9, atol=1e-12 * np.abs(duo).max())
    np.testing.assert_allclose(dq, dqo, rtol=1e-9, atol=1e-12 * np.abs(dqo).max())
    assert np.abs(dq).max() > 0


def test_forward_sweep_is_the_production_step(box, device_grad):
    core = device_grad[3]
    sol = GpuSolver(box.topo, box.params(64))
    sol.set_bc_kinds(box.kinds)
    sol.set_links(box.link, box.q)
    sol.init_equilibrium(1.0, box.init_u0, box.init_mask)
    sol.advance(STEPS)
    ref = sol.download_core()
    sol.close()
    np.testing.assert_array_equal(core, ref.reshape(core.shape))


def test_checkpoint_interval_does_not_change_results(box, device_grad):
    for k in (1, STEPS):
        print(f"checkpoint {k}", flush=True)
        J, du, dq = DragAdjoint(box, STEPS, WIN, checkpoint=k).evaluate()
        assert J == device_grad[0]
        np.testing.assert_array_equal(du, device_grad[1])
        np.testing.assert_array_equal(dq, device_grad[2])


def test_central_differences_of_the_device_objective(box, device_grad):
    adj = DragAdjoint(box, STEPS, WIN, checkpoint=10)
    eps = 1e-6
    u = np.array(box.init_u0, float)
    up, um = u.copy(), u.copy()
    up[0] += eps
    um[0] -= eps
    fd = (adj.evaluate(u0=up)[0] - adj.evaluate(u0=um)[0]) / (2 * eps)
    assert device_grad[1][0] == pytest.approx(fd, rel=1e-6)
    k = int(np.argmax(np.abs(device_grad[2])))
    qp, qm = box.q.copy(), box.q.copy()
    qp[k] += eps
    qm[k] -= eps
    fd = (adj.evaluate(q=qp)[0] - adj.evaluate(q=qm)[0]) / (2 * eps)
    assert device_grad[2][k] == pytest.approx(fd, rel=1e-4)


def test_autograd_function(box, device_grad):
    adj = DragAdjoint(box, STEPS, WIN, checkpoint=7)
    u0 = torch.tensor(box.init_u0, dtype=torch.float64, requires_grad=True)
    q = torch.tensor(box.q, dtype=torch.float64, requires_grad=True)
    J = DragObjective.apply(u0, q, adj)
    (2.0 * J).backward()
    assert float(J.detach()) == device_grad[0]
    np.testing.assert_array_equal(u0.grad.numpy(), 2.0 * device_grad[1])
    np.testing.assert_array_equal(q.grad.numpy(), 2.0 * device_grad[2])


def test_create_scope():
    # velocity / pressure boundaries and several levels run the multi-level
    # engine; ConstRhoBGK and ensembles are not differentiated
    open_channel = cases.cylinder_channel(bricks_x=4, base_level=0, fine=False, bricks_yz=2, xc=0.6,
                                          radius=0.125, u_in=0.05, tau=0.6)
    sol = GpuSolver(open_channel.topo, open_channel.params(64))
    sol.set_bc_kinds(open_channel.kinds)
    h = C.c_void_p()
    assert abi.lib().octlb_adj_create(sol._h, C.byref(h)) == 0 and h.value
    abi.lib().octlb_adj_destroy(h)
    sol.close()
    ens = cases.cavity_ensemble(count=2, sizes=(32,), seed=2)
    e = GpuSolver.ensemble(ens, 64)
    e.set_bc_kinds(ens.kinds)
    h = C.c_void_p()
    rc = abi.lib().octlb_adj_create(e._h, C.byref(h))
    msg = abi.lib().octlb_last_error().decode()
    e.close()
    assert rc == -1 and "ensembles" in msg  # OCTLB_EINVAL


def test_fp32_forward_fp64_adjoint(box, device_grad):
    J, du, dq, core = DragAdjoint(box, STEPS, WIN, checkpoint=7, precision=32).evaluate(final_core=True)
    sol = GpuSolver(box.topo, box.params(32))
    sol.set_bc_kinds(box.kinds)
    sol.set_links(box.link, box.q)
    sol.init_equilibrium(1.0, box.init_u0, box.init_mask)
    sol.advance(STEPS)
    ref = sol.download_core()
    sol.close()
    np.testing.assert_array_equal(core, ref.reshape(core.shape))
    J64, du64, dq64, _ = device_grad
    rel_u = abs(du[0] - du64[0]) / abs(du64[0])
    rel_q = np.abs(dq - dq64).max() / np.abs(dq64).max()
    print(f"FP32 forward vs FP64: J {abs(J - J64) / abs(J64):.2e}, dJ/du0 {rel_u:.2e}, dJ/dq {rel_q:.2e}", flush=True)
    assert abs(J - J64) <= 1e-4 * abs(J64)
    assert rel_u <= 1e-4 and rel_q <= 1e-4
