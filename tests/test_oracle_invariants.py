"""Known answers and invariants that pin the parts of the oracle the
reference harness cannot run here (multi-level coupling and the recursive
subcycle need the reference's p4est-backed LevelCoupler / TimeLoop):

* TimeLoop counts 1:2:4 per coarse step and the coupler call order
  (tests/unit/solver/test_time_loop_levels.cpp:66-149).
* Coupling plan: unique keys, indices in range, coarse side on the interface
  layer, every fine slot present (test_lagrava_coupler.cpp:126-192).
* Prolongation at rest / after full time: fine rho,u equal coarse rho,u
  (:194-225); prolong + restrict conserve mass at rest (:227-248).
* Momentum-exchange drag of one link at rest = -1/9 (test_drag_postprocessor.cpp:29-47).
* Bouzidi pull with q = 0.35 (test_bouzidi_link.cpp:105-136) and the two
  branches of bouzidi_pull.h at q = 1/2 (both reduce to half-way bounce-back).
* Equilibrium moments: sum feq = rho - 1 and sum c feq = rho u (shifted storage).
Each check is computed independently of the oracle's formulas (closed forms,
brute-force index arithmetic), so a dropped term or transposed index fails.
"""
import numpy as np
import pytest

from octlb_amd import abi, cases
from octlb_amd.forest import Forest, single_block_topology
from oracle.binding import OracleSim

C = cases.C
OPP = cases.OPP
W = np.array([1 / 3] + [1 / 18] * 3 + [1 / 36] * 6 + [1 / 18] * 3 + [1 / 36] * 6)


def _noop_case(topo, omega=1.0):
    kinds = np.zeros((topo.num_blocks, topo.n, topo.n, topo.n), np.uint8)
    return cases.Case(name="noop", topo=topo, omega=omega, handler=abi.HANDLER_NOOP, const_rho=0,
                      bouzidi=1, snapshot=0, specs=[], kinds=kinds)


def _three_level_forest():
    # MakeThreeLevelForest (test_time_loop_levels.cpp:44-52): level 1 everywhere,
    # the centre octant refined once more, then balanced
    f = Forest([0, 0, 0, 1, 1, 1])
    f.refine_uniform(1)
    f.refine_box([0.1, 0.1, 0.1, 0.2, 0.2, 0.2], 2)
    f.balance()
    return f.topology(4, 1)


def test_subcycle_counts_and_call_order():
    topo = _three_level_forest()
    assert set(topo.level) == {1, 2}
    sim = OracleSim(topo, _noop_case(topo).params())
    sim.advance(1)
    col, st, calls = sim.counters()
    # levels 0 (empty), 1, 2 run 1 : 2 : 4 times per coarse step
    assert list(col[:3]) == [1, 2, 4] and list(st[:3]) == [1, 2, 4]
    expected = []

    def rec(level, max_level):
        if level >= max_level:
            return
        expected.append((0, level))
        rec(level + 1, max_level)
        expected.append((1, level))
        rec(level + 1, max_level)
        expected.append((2, level))

    rec(0, 2)
    assert [tuple(c) for c in calls] == expected


def test_coupling_plan_properties():
    f = Forest([0, 0, 0, 1, 1, 1])
    f.refine_uniform(1)
    f.refine_box([0.1, 0.1, 0.1, 0.2, 0.2, 0.2], 2)
    f.balance()
    n = 4
    topo = f.topology(n, 1)
    sim = OracleSim(topo, _noop_case(topo).params())
    plan = sim.plan()
    assert len(plan) == topo.num_cf * n * n  # 4 slots x (n/2)^2 points per face
    keys = {tuple(r[:8]) for r in plan}
    assert len(keys) == len(plan)
    assert plan[:, [1, 2, 3, 5, 6, 7]].min() >= 0 and plan[:, [1, 2, 3, 5, 6, 7]].max() < n
    for r in plan:
        axis, is_max = r[8] // 2, r[8] & 1
        assert r[1 + axis] == (n - 1 if is_max else 0)        # coarse interface layer
        assert r[5 + axis] == (0 if is_max else n - 1)        # fine interface layer
    for k in range(topo.num_cf):
        cf = topo.coarse_fine[k]
        for slot in range(4):
            rows = plan[(plan[:, 0] == cf.coarse_id) & (plan[:, 4] == cf.fine_ids[slot])]
            assert len(rows) == (n // 2) ** 2 and set(rows[:, 11]) == {slot}
    # the quarter a fine block couples to is the one its centre lies over
    b = topo.bounds
    for r in plan:
        cc = 0.5 * (b[r[0], :3] + b[r[0], 3:])
        fc = 0.5 * (b[r[4], :3] + b[r[4], 3:])
        axis = r[8] // 2
        for t in ((axis + 1) % 3, (axis + 2) % 3):
            half = (r[1 + t] >= n // 2)
            assert half == (fc[t] >= cc[t])
            assert r[5 + t] == r[1 + t] - (n // 2 if half else 0)


def _uniform_state_case():
    f = Forest([0, 0, 0, 1, 1, 1])
    f.refine_uniform(1)
    f.refine_box([0.1, 0.1, 0.1, 0.2, 0.2, 0.2], 2)
    f.balance()
    topo = f.topology(4, 1)
    return topo


def test_rest_state_is_a_fixed_point_and_conserves_mass():
    topo = _uniform_state_case()
    case = _noop_case(topo)
    sim = OracleSim(topo, case.params())
    sim.set_bc_kinds(case.kinds)
    sim.init_equilibrium(1.0, (0.0, 0.0, 0.0), 0xFF)
    m0 = sim.mass()
    sim.advance(3)
    core = sim.download_core()
    assert np.abs(core).max() == 0.0          # shifted rest equilibrium is all zeros
    assert abs(sim.mass() - m0) <= 1e-10


def test_full_time_prolongation_copies_coarse_macros():
    """After a coarse step from a uniform moving state every prolonged fine
    cell carries exactly the coarse cell's rho and u (scaling only touches
    f_neq, which is zero at equilibrium)."""
    topo = _uniform_state_case()
    case = _noop_case(topo)
    sim = OracleSim(topo, case.params())
    sim.set_bc_kinds(case.kinds)
    u0 = (0.02, -0.01, 0.005)
    sim.init_equilibrium(1.01, u0, 0xFF)
    plan = sim.plan()
    core = sim.download_core()

    def macro(f):
        rho = 1.0 + f.sum()
        return rho, (f[:, None] * C).sum(0) / rho

    for r in plan[:20]:
        rc, uc = macro(core[r[0], r[1], r[2], r[3]])
        rf, uf = macro(core[r[4], r[5], r[6], r[7]])
        assert abs(rc - rf) < 1e-14 and np.abs(uc - uf).max() < 1e-14
        assert abs(rc - 1.01) < 1e-14 and np.abs(uc - np.array(u0)).max() < 1e-14


def test_drag_single_link_known_force():
    """One bulk cell at f_i = t_i with its only solid neighbour at +x gives
    force_on_fluid = -(t_10 + t_1) e_x = -1/9 e_x."""
    n = 4
    topo = single_block_topology(n, 1)
    kinds = np.zeros((1, n, n, n), np.uint8)
    kinds[0, 1, 0, 0] = abi.KIND_SOLID
    case = cases.Case(name="drag", topo=topo, omega=1.0, handler=abi.HANDLER_NOOP, const_rho=0,
                      bouzidi=1, snapshot=0, specs=[], kinds=kinds)
    sim = OracleSim(topo, case.params())
    sim.set_bc_kinds(kinds)
    blk = np.zeros((n + 2, n + 2, n + 2, 19))
    blk[1, 1, 1, :] = W
    sim.upload_block(0, blk)
    f = sim.drag(np.array([-10, -10, -10, 10, 10, 10.0]))
    assert abs(f[0] + 1.0 / 9.0) < 1e-12 and abs(f[1]) < 1e-12 and abs(f[2]) < 1e-12


def _bouzidi(fbb, fint, fsame, q):
    # closed form of the reference's two branches (bouzidi_pull.h:8-26)
    if q >= 0.5:
        return fbb / (2 * q) + (1 - 1 / (2 * q)) * fint
    return 2 * q * fbb + (1 - 2 * q) * fsame


@pytest.mark.parametrize("q", [0.35, 0.5, 0.8])
def test_bouzidi_pull_known_answer(q):
    """Stream_BouzidiReplacesSolidPull: a kBulk cell whose +x neighbour is
    solid pulls population 1 (c = -x) through the Bouzidi formula."""
    n = 8
    topo = single_block_topology(n, 1)
    kinds = np.zeros((1, n, n, n), np.uint8)
    kinds[0, n // 2:, :, :] = abi.KIND_SOLID
    i = n // 2 - 1
    ipop = 1
    case = cases.Case(name="bz", topo=topo, omega=1.0, handler=abi.HANDLER_NOOP, const_rho=0,
                      bouzidi=1, snapshot=0, specs=[], kinds=kinds,
                      link=np.array([[0, (i * n + 3) * n + 3, ipop]], np.int32), q=np.array([q]))
    sim = OracleSim(topo, case.params())
    sim.set_bc_kinds(kinds)
    sim.set_links(case.link, case.q)
    rng = np.random.default_rng(3)
    blk = rng.normal(0, 0.01, (n + 2, n + 2, n + 2, 19))
    sim.upload_block(0, blk)
    sim.block_collide(0)
    post = sim.download_block(0)
    sim.block_stream(0)
    got = sim.download_block(0)[i + 1, 4, 4, ipop]
    cell = post[i + 1, 4, 4]
    nb = post[i + 2, 4, 4]  # x + c_1 = x - e_x ... neighbour at x + c_i
    nb = post[i + 1 - 1, 4, 4]
    exp = _bouzidi(cell[OPP[ipop]], nb[ipop], cell[ipop], q)
    assert abs(got - exp) < 1e-12
    if q == 0.5:
        assert abs(got - cell[OPP[ipop]]) < 1e-15


def test_equilibrium_moments():
    """feq written by init: sum = rho - 1 and first moment = rho u."""
    n = 4
    topo = single_block_topology(n, 1)
    case = _noop_case(topo)
    sim = OracleSim(topo, case.params())
    sim.set_bc_kinds(case.kinds)
    rho, u = 1.03, np.array([0.04, -0.02, 0.01])
    sim.init_equilibrium(rho, u, 0xFF)
    f = sim.download_core()[0, 1, 2, 3]
    assert abs(f.sum() - (rho - 1.0)) < 1e-15
    assert np.abs((f[:, None] * C).sum(0) - rho * u).max() < 1e-15
    # second moment: sum c_a c_b (f + t) = rho u_a u_b + rho/3 delta
    P = ((f + W)[:, None, None] * C[:, :, None] * C[:, None, :]).sum(0)
    assert np.abs(P - (rho * np.outer(u, u) + rho / 3 * np.eye(3))).max() < 1e-15


def test_edge_exchange_requires_same_level_face_path():
    topo = _uniform_state_case()
    sim = OracleSim(topo, _noop_case(topo).params())
    edges = sim.edges()
    same = {(a, d): r for a, d, r in topo.same_level}
    for a, d1, d2, dd in edges:
        b = same[(a, d1)]
        assert same[(b, d2)] == dd
        assert topo.level[a] == topo.level[dd]
        # dd is the diagonal neighbour: centres differ by one block along both axes
        ca = 0.5 * (topo.bounds[a, :3] + topo.bounds[a, 3:])
        cd = 0.5 * (topo.bounds[dd, :3] + topo.bounds[dd, 3:])
        size = topo.bounds[a, 3] - topo.bounds[a, 0]
        delta = np.round((cd - ca) / size).astype(int)
        exp = np.zeros(3, int)
        exp[d1 // 2] = 1 if d1 & 1 else -1
        exp[d2 // 2] = 1 if d2 & 1 else -1
        assert (delta == exp).all()
