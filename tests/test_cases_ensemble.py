"""Ensemble builder: member cases concatenate into one block list whose
faces never join two members, whose member ranges cover the blocks, and
whose per-member topology (after removing the block offset) is the member's
own.  The device side is checked against single-case runs in
test_gpu_ensemble.py."""
import numpy as np

from octlb_amd import abi, cases


def test_cavity_blocked_walls_and_lid():
    c = cases.cavity_blocked(16, tau=0.6, u_lid=0.05)
    assert c.topo.num_blocks == 8 and c.const_rho == 1
    k = c.kinds.reshape(2, 2, 2, 8, 8, 8)
    assert (c.kinds == abi.KIND_VELOCITY_DIRICHLET).sum() == 16 ** 3 - 14 ** 3
    lid = [s for s in c.specs if s.face == abi.YMAX][0]
    assert lid.u_wall[0] == 0.05 and k.shape[0] == 2


def test_ensemble_concatenation():
    ms = [cases.cavity_blocked(16, 0.6, 0.05), cases.cavity_blocked(24, 0.55, 0.03, x0=2.0),
          cases.cavity_blocked(16, 0.7, 0.08, x0=4.0)]
    e = cases.ensemble(ms)
    assert list(e.member_begin) == [0, 8, 35, 43]
    owner = np.repeat(np.arange(3), np.diff(e.member_begin))
    s = e.topo.same_level
    assert np.all(owner[s[:, 0]] == owner[s[:, 2]])
    for m, c in enumerate(ms):
        lo = e.member_begin[m]
        mine = s[owner[s[:, 0]] == m].copy()
        mine[:, [0, 2]] -= lo
        assert sorted(map(tuple, mine)) == sorted(map(tuple, c.topo.same_level))
        np.testing.assert_array_equal(e.kinds[lo:e.member_begin[m + 1]], c.kinds)
    p = e.params_for(32)
    assert [q.omega for q in p] == [1 / 0.6, 1 / 0.55, 1 / 0.7] and all(q.precision == 32 for q in p)


def test_cavity_ensemble_config():
    e = cases.cavity_ensemble(count=8, seed=3)
    sizes = [c.info["cells"] for c in e.members]
    assert set(sizes) <= {32, 48, 64, 96}
    assert all(0.52 <= c.info["tau"] <= 0.8 and 0.02 <= c.info["u_lid"] <= 0.08 for c in e.members)
    assert e.topo.num_blocks == sum((s // 8) ** 3 for s in sizes)
