"""Ensembles (octlb_dev_create_ensemble): independent cavities of different
sizes, relaxation rates and lid speeds stepped as one device state.  Every
member must evolve exactly as the same case run alone: FP64 members
bit-equal to the oracle's single-case runs (populations and ConstRho
average), FP32 members bit-equal to single-case GPU runs (eager and graph
replay), and a block face joining two members is refused.
"""
import ctypes as C

import numpy as np
import pytest

from octlb_amd import abi, cases
from octlb_amd.solver import GpuSolver
from tests.util import make_gpu, make_oracle

pytestmark = pytest.mark.gpu

MEMBERS = [(16, 0.6, 0.05), (24, 0.55, 0.03), (16, 0.7, 0.08)]
STEPS = 12


def _ensemble(fast_stats=0):
    ms = [cases.cavity_blocked(n, t, u, x0=2.0 * i) for i, (n, t, u) in enumerate(MEMBERS)]
    return ms, cases.ensemble(ms, fast_stats=fast_stats)


def _start(e, precision):
    g = GpuSolver.ensemble(e, precision)
    g.set_bc_kinds(e.kinds)
    g.init_equilibrium(1.0, (0.0, 0.0, 0.0), 0xFF)
    return g


def test_fp64_members_match_oracle_bitwise():
    ms, e = _ensemble()
    g = _start(e, 64)
    g.advance(STEPS)
    core = g.download_core()
    for m, c in enumerate(ms):
        print(f"member {m}: {c.info}", flush=True)
        o = make_oracle(c)
        o.advance(STEPS)
        lo, hi = e.member_begin[m], e.member_begin[m + 1]
        np.testing.assert_array_equal(core[lo:hi], o.download_core())
        assert g.member_average_rho(m) == o.average_rho()
    assert g.average_rho() == g.member_average_rho(0)


@pytest.mark.parametrize("graph", [False, True])
def test_fp32_members_match_single_runs(graph):
    ms, e = _ensemble()
    g = _start(e, 32)
    (g.advance_graph if graph else g.advance)(STEPS + 7)
    core = g.download_core()
    for m, c in enumerate(ms):
        s = make_gpu(c, precision=32)
        (s.advance_graph if graph else s.advance)(STEPS + 7)
        lo, hi = e.member_begin[m], e.member_begin[m + 1]
        np.testing.assert_array_equal(core[lo:hi], s.download_core())
        assert g.member_average_rho(m) == s.average_rho()


def test_fast_stats_members_match_single_runs():
    ms, e = _ensemble(fast_stats=1)
    g = _start(e, 32)
    g.advance(STEPS)
    core = g.download_core()
    for m, c in enumerate(ms):
        p = c.params(32)
        p.fast_stats = 1
        s = GpuSolver(c.topo, p)
        s.set_bc_kinds(c.kinds)
        s.init_equilibrium(1.0, (0.0, 0.0, 0.0), 0xFF)
        s.advance(STEPS)
        lo, hi = e.member_begin[m], e.member_begin[m + 1]
        np.testing.assert_array_equal(core[lo:hi], s.download_core())


def test_face_between_members_is_refused():
    c = cases.cavity_blocked(16)
    e = cases.ensemble([c])
    e.member_begin = np.array([0, 4, 8], np.int32)  # splits a connected cavity
    e.members = [c, c]
    with pytest.raises(abi.OctlbError, match="face joins two ensemble members"):
        GpuSolver.ensemble(e, 64)
