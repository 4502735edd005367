"""Pins the oracle's Lagrava coupler (prolongation at half and full time,
prev refresh, restriction with the in-block neighbour average) against an
independent numpy statement of level_coupler.cpp:163-214, 438-490, 613-654 on
random non-equilibrium states.  The reference's own LevelCoupler needs the
p4est-backed forest and cannot run here, so these checks, the plan
properties and the rest-state invariants (test_oracle_invariants.py) are
what pin the coupler.

A plausible slip -- feq of the averaged macros replaced by the average of
feqs, a wrong scale (s_p = (tau - 1/4)/tau, s_r = tau/(tau - 1/4)), a wrong
quadrant or fine index, a neighbour set other than the D3Q19 neighbours
inside the fine block, prev not refreshed at full time -- fails one of them.
"""
import numpy as np

from octlb_amd import abi, cases
from octlb_amd.forest import Forest
from oracle.binding import OracleSim

C = cases.C.astype(np.float64)
W = np.array([1 / 3] + [1 / 18] * 3 + [1 / 36] * 6 + [1 / 18] * 3 + [1 / 36] * 6)
OMEGA = 1.0 / 0.57
TAU = 1.0 / OMEGA
S_P = (TAU - 0.25) / TAU
S_R = TAU / (TAU - 0.25)


def _sim():
    f = Forest([0, 0, 0, 1, 1, 1])
    f.refine_uniform(1)
    f.refine_box([0.1, 0.1, 0.1, 0.2, 0.2, 0.2], 2)
    f.balance()
    topo = f.topology(4, 1)
    kinds = np.zeros((topo.num_blocks, 4, 4, 4), np.uint8)
    case = cases.Case(name="couple", topo=topo, omega=OMEGA, handler=abi.HANDLER_NOOP,
                      const_rho=0, bouzidi=1, snapshot=0, specs=[], kinds=kinds)
    sim = OracleSim(topo, case.params())
    sim.set_bc_kinds(kinds)
    return sim, topo


def _upload_random(sim, topo, seed):
    rng = np.random.default_rng(seed)
    NP = topo.n + 2
    blocks = rng.normal(0.0, 0.01, (topo.num_blocks, NP, NP, NP, 19))
    for b in range(topo.num_blocks):
        sim.upload_block(b, blocks[b])
    return sim.download_core()


def _macro(f):
    rho = 1.0 + f.sum()
    u = (f[:, None] * C).sum(0) / rho
    return rho, u, f - _feq(rho, u)


def _feq(rho, u):
    cu = C @ u
    return W * rho * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * (u @ u)) - W


def _check(got, exp, tol=1e-13):
    assert np.abs(got - exp).max() <= tol * max(np.abs(exp).max(), 1e-3)


def _prolong_expected(core, plan, level, prev=None):
    exp = core.copy()
    for r in plan:
        if r[9] != level:
            continue
        rho, u, fneq = _macro(core[r[0], r[1], r[2], r[3]])
        if prev is not None:
            prho, pu, pfneq = prev[tuple(r[:4])]
            rho, u, fneq = 0.5 * (prho + rho), 0.5 * (pu + u), 0.5 * (pfneq + fneq)
        exp[r[4], r[5], r[6], r[7]] = _feq(rho, u) + S_P * fneq
    return exp


def test_full_time_prolongation():
    sim, topo = _sim()
    core = _upload_random(sim, topo, 1)
    plan = sim.plan()
    assert (plan[:, 9] == 1).any()
    sim.apply_prolongation(1, half_time=False)
    _check(sim.download_core(), _prolong_expected(core, plan, 1))


def test_half_time_prolongation_mixes_prev_and_current():
    sim, topo = _sim()
    core_a = _upload_random(sim, topo, 2)
    sim.coupler_init()
    plan = sim.plan()
    prev = {tuple(r[:4]): _macro(core_a[r[0], r[1], r[2], r[3]]) for r in plan}
    core_b = _upload_random(sim, topo, 3)
    sim.apply_prolongation(1, half_time=True)
    _check(sim.download_core(), _prolong_expected(core_b, plan, 1, prev))


def test_full_time_refreshes_prev():
    sim, topo = _sim()
    _upload_random(sim, topo, 4)
    sim.coupler_init()
    core_b = _upload_random(sim, topo, 5)
    sim.apply_prolongation(1, half_time=False)
    plan = sim.plan()
    prev = {tuple(r[:4]): _macro(core_b[r[0], r[1], r[2], r[3]]) for r in plan if r[9] == 1}
    core_c = _upload_random(sim, topo, 6)
    sim.apply_prolongation(1, half_time=True)
    _check(sim.download_core(), _prolong_expected(core_c, plan, 1, prev))


def test_restriction_in_block_neighbour_average():
    sim, topo = _sim()
    core = _upload_random(sim, topo, 7)
    plan = sim.plan()
    n = topo.n
    sim.apply_restriction(2)
    exp = core.copy()
    for r in plan:
        if r[10] != 2:
            continue
        fb, x, y, z = r[4], r[5], r[6], r[7]
        rho, u, fneq = _macro(core[fb, x, y, z])
        acc, cnt = fneq.copy(), 1
        for j in range(1, 19):
            p = np.array([x, y, z]) + cases.C[j]
            if (p < 0).any() or (p >= n).any():
                continue
            acc += _macro(core[fb, p[0], p[1], p[2]])[2]
            cnt += 1
        exp[r[0], r[1], r[2], r[3]] = _feq(rho, u) + S_R * acc / cnt
    _check(sim.download_core(), exp)


def test_restriction_of_equilibrium_is_equilibrium():
    """fneq = 0 everywhere -> the coarse cell gets feq(rho_f, u_f) exactly,
    whatever the averaging set."""
    sim, topo = _sim()
    sim.init_equilibrium(1.02, (0.03, -0.01, 0.02), 0xFF)
    core = sim.download_core()
    sim.apply_restriction(2)
    _check(sim.download_core(), core, 1e-14)
