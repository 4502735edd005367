"""octlb_dev_migrate: the state of one forest moved onto the forest after a
refinement and back.  Unchanged blocks are copied bit for bit; refined
blocks carry their parent cells' olb (rho, u) and s_p times the
non-equilibrium part; refine-then-coarsen returns the original populations
to rounding (s_r s_p = 1, every child of a cell identical); mass in
fine-equivalent units is conserved.  The reference has no migration code,
so the interpolation rules are checked against their own statement.
"""
import numpy as np
import pytest

from octlb_amd import cases
from octlb_amd.forest import Forest
from octlb_amd.solver import GpuSolver

pytestmark = pytest.mark.gpu

C = cases.C
TAU = 0.6


def _topo(refine, n=8):
    f = Forest(np.array([0.0, 0.0, 0.0, 2.0, 1.0, 1.0]), (2, 1, 1))
    f.refine_uniform(1)
    if refine:
        f.refine_box([0.01, 0.01, 0.01, 0.49, 0.49, 0.49], 2)
        f.balance()
    return f.topology(n, 1)


def _solver(topo, precision=64):
    from octlb_amd import abi
    s = GpuSolver(topo, abi.make_params(1.0 / TAU, precision=precision))
    s.set_bc_kinds(np.zeros((topo.num_blocks, topo.n, topo.n, topo.n), np.uint8))
    return s


def _tasks(old, new):
    """(mode, old blocks, octant) of every block of `new`: the same block,
    the parent (octant of the child in it) or the eight children."""
    out = []
    for j in range(new.num_blocks):
        bb, lv = new.bounds[j], new.level[j]
        same = [i for i in range(old.num_blocks) if old.level[i] == lv and np.all(old.bounds[i] == bb)]
        if same:
            out.append((0, same, (0, 0, 0)))
            continue
        par = [i for i in range(old.num_blocks) if old.level[i] == lv - 1 and np.all(old.bounds[i][:3] <= bb[:3])
               and np.all(old.bounds[i][3:] >= bb[3:])]
        if par:
            pb = old.bounds[par[0]]
            oct = tuple(int(v) for v in np.rint((bb[:3] - pb[:3]) / (bb[3] - bb[0])))
            out.append((1, par, oct))
            continue
        w = (bb[3] - bb[0]) / 2
        kids = []
        for c in range(8):
            lo = bb[:3] + w * np.array([c & 1, (c >> 1) & 1, c >> 2])
            kids += [i for i in range(old.num_blocks) if old.level[i] == lv + 1 and np.allclose(old.bounds[i][:3], lo)]
        assert len(kids) == 8
        out.append((2, kids, (0, 0, 0)))
    return out


@pytest.mark.parametrize("precision,n", [(64, 8), (32, 8)])
def test_device_migration_matches_the_oracle(precision, n):
    # every new block's core cells, refine and then coarsen, against the
    # oracle's statement of the rules fed with the old handle's downloaded
    # halo-inclusive views: bit for bit in both precisions
    from oracle import binding
    ta, tb = _topo(False, n), _topo(True, n)
    src = _solver(ta, precision)
    nc = ta.num_blocks * n ** 3
    rng = np.random.default_rng(11)
    src.init_field(1.0 + 0.01 * rng.random(nc), 0.02 * (rng.random(nc * 3) - 0.5))
    src.advance(2)
    dst = _solver(tb, precision)
    back = _solver(ta, precision)
    modes = set()
    for old, new, a, b in ((ta, tb, src, dst), (tb, ta, dst, back)):
        print(f"migrate precision {precision} n {n}: {old.num_blocks} -> {new.num_blocks} blocks", flush=True)
        b.migrate_from(a)
        got = b.download_core()
        views = [a.download_block(i) for i in range(old.num_blocks)]
        for j, (mode, srcs, oct) in enumerate(_tasks(old, new)):
            modes.add(mode)
            want = binding.migrate_block(mode, n, 1, oct, [views[i] for i in srcs], TAU)
            want = want.reshape(n + 2, n + 2, n + 2, 19)[1:-1, 1:-1, 1:-1]
            if precision == 32:
                want = want.astype(np.float32).astype(np.float64)
            np.testing.assert_array_equal(got[j], want, err_msg=f"block {j} mode {mode}")
    assert modes == {0, 1, 2}
    for s in (src, dst, back):
        s.close()


def _macro(f):
    rho = f.sum(-1) + 1.0
    u = np.einsum("...i,id->...d", f, C.astype(float)) / rho[..., None]
    usqr = (u * u).sum(-1)
    w = np.array([1 / 3] + [1 / 18] * 3 + [1 / 36] * 6 + [1 / 18] * 3 + [1 / 36] * 6)
    cu = np.einsum("...d,id->...i", u, C.astype(float))
    feq = rho[..., None] * w * (1 + 3 * cu + 4.5 * cu * cu - 1.5 * usqr[..., None]) - w
    return rho, u, f - feq


def test_refine_and_coarsen_round_trip():
    ta, tb = _topo(False), _topo(True)
    assert tb.num_blocks == ta.num_blocks + 7
    src = _solver(ta)
    n = 8
    nc = ta.num_blocks * n ** 3
    rng = np.random.default_rng(3)
    src.init_field(1.0 + 0.01 * rng.random(nc), 0.02 * (rng.random(nc * 3) - 0.5))
    src.advance(3)  # a non-equilibrium part
    a = src.download_core()
    dst = _solver(tb)
    dst.migrate_from(src)
    b = dst.download_core()
    tau_p = (TAU - 0.25) / TAU
    refined = 0
    for j in range(tb.num_blocks):
        bb = tb.bounds[j]
        same = [i for i in range(ta.num_blocks) if ta.level[i] == tb.level[j] and np.all(ta.bounds[i] == bb)]
        if same:
            np.testing.assert_array_equal(b[j], a[same[0]])
            continue
        par = [i for i in range(ta.num_blocks) if ta.level[i] == tb.level[j] - 1 and np.all(ta.bounds[i][:3] <= bb[:3])
               and np.all(ta.bounds[i][3:] >= bb[3:])][0]
        refined += 1
        pb = ta.bounds[par]
        off = np.rint((bb[:3] - pb[:3]) / ((pb[3] - pb[0]) / n)).astype(int)
        idx = (np.arange(n) // 2)
        parent_cells = a[par][np.ix_(off[0] + idx, off[1] + idx, off[2] + idx)]
        rp, up, fp = _macro(parent_cells)
        rc, uc, fc = _macro(b[j])
        np.testing.assert_allclose(rc, rp, rtol=1e-14)
        np.testing.assert_allclose(uc, up, rtol=0, atol=1e-15)
        np.testing.assert_allclose(fc, tau_p * fp, rtol=0, atol=1e-15)
    assert refined == 8
    back = _solver(ta)
    back.migrate_from(dst)
    c = back.download_core()
    np.testing.assert_allclose(c, a, rtol=0, atol=1e-14)
    # fine-equivalent mass: a level-2 cell is 1/8 of a level-1 cell
    ma = (a.sum(-1) + 1.0).sum()
    wb = np.where(tb.level == 2, 1.0 / 8.0, 1.0)
    mb = ((b.sum(-1) + 1.0).sum(axis=(1, 2, 3)) * wb).sum()
    assert mb == pytest.approx(ma, rel=1e-13)
    for s in (src, dst, back):
        s.close()


def test_identity_and_mismatch():
    from octlb_amd import abi
    ta = _topo(False)
    src = _solver(ta)
    src.init_equilibrium(1.0, (0.03, 0.0, 0.01), 0xFF)
    src.advance(2)
    dst = _solver(ta)
    dst.migrate_from(src)
    np.testing.assert_array_equal(dst.download_core(), src.download_core())
    other = GpuSolver(ta, abi.make_params(1.0 / 0.7, precision=64))
    with pytest.raises(abi.OctlbError):
        other.migrate_from(src)
    for s in (src, dst, other):
        s.close()


def test_migrated_refined_state_steps_like_the_oracle():
    # the first steps after a migration onto the refined forest: the GPU and
    # the oracle, both started from the migrated halo-inclusive views,
    # agree to 1e-12 (new blocks' halo slots hold defined values)
    from octlb_amd import abi
    from oracle.binding import OracleSim
    ta, tb = _topo(False), _topo(True)
    src = _solver(ta)
    nc = ta.num_blocks * 8 ** 3
    rng = np.random.default_rng(5)
    src.init_field(1.0 + 0.01 * rng.random(nc), 0.02 * (rng.random(nc * 3) - 0.5))
    src.advance(3)
    dst = _solver(tb)
    dst.migrate_from(src)
    ora = OracleSim(tb, abi.make_params(1.0 / TAU, precision=64))
    ora.set_bc_kinds(np.zeros((tb.num_blocks, 8, 8, 8), np.uint8))
    for b in range(tb.num_blocks):
        ora.upload_block(b, dst.download_block(b))
    dst.advance(4)
    ora.advance(4)
    g, o = dst.download_core(), ora.download_core()
    assert np.abs(g - o).max() <= 1e-12 * np.abs(o).max()
    for s in (src, dst, ora):
        s.close()


def test_migrate_carries_step_and_const_rho_average():
    # a ramped inlet and ConstRhoBGK: the step counter and the lagged average
    # move with the populations, so the migrated handle continues bit for bit
    import dataclasses
    case = cases.cylinder_channel(bricks_x=2, base_level=1, fine=True, xc=0.5, radius=0.1,
                                  refine_lo=0.25, refine_hi=1.0, ramp_end_t=40.0)
    case = dataclasses.replace(case, const_rho=1)
    src = case.apply(GpuSolver(case.topo, case.params(64)))
    src.advance(5)
    dst = case.apply(GpuSolver(case.topo, case.params(64)))
    dst.migrate_from(src)
    assert dst.average_rho() == src.average_rho()
    src.advance(6)
    dst.advance(6)
    np.testing.assert_array_equal(dst.download_core(), src.download_core())
    np.testing.assert_array_equal(dst.drag(case.drag_bbox), src.drag(case.drag_bbox))
    assert dst.average_rho() == src.average_rho()
    for s in (src, dst):
        s.close()


def test_migrate_rejects_ensembles():
    from octlb_amd import abi
    ens = cases.cavity_ensemble(count=2, sizes=(32,), seed=1)
    a = GpuSolver.ensemble(ens, 64)
    b = GpuSolver.ensemble(ens, 64)
    with pytest.raises(abi.OctlbError, match="ensemble"):
        b.migrate_from(a)
    for s in (a, b):
        s.close()


def test_padded_block_copies_whole_view():
    # the padded layout (one block, any edge) keeps every halo slot: a
    # migration onto the same forest copies the whole view
    case = cases.cavity(16, tau=TAU, u_lid=0.05)
    a = case.apply(GpuSolver(case.topo, case.params(64)))
    a.advance(3)
    b = case.apply(GpuSolver(case.topo, case.params(64)))
    b.migrate_from(a)
    np.testing.assert_array_equal(b.download_block(0), a.download_block(0))
    for s in (a, b):
        s.close()
