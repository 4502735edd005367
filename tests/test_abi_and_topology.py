"""C-ABI library loading / exports and the brick-forest topology.

The forest replaces p4est (absent here).  p4est's exact octant and face
enumeration order is not reproducible without it, so the tables are pinned
as canonical sets against an independent brute-force enumeration by
geometry, and against the counts and properties the reference's own mesh
tests assert (tests/unit/mesh/test_octree_forest.cpp:152-214,
test_face_pair_list.cpp:47-125).
"""
import itertools
import re
from pathlib import Path

import numpy as np
import pytest

from octlb_amd import abi
from octlb_amd.forest import Forest

ROOT = Path(__file__).resolve().parent.parent


def test_library_exports_every_declared_symbol():
    header = (ROOT / "include" / "octlb_gpu.h").read_text()
    declared = set(re.findall(r"\b(octlb_[a-z_]+)\s*\(", header))
    assert declared, "no declarations parsed"
    lib = abi.lib()
    missing = [s for s in sorted(declared) if not hasattr(lib, s)]
    assert not missing, f"undefined exports: {missing}"
    assert declared <= set(abi.EXPORTED) | {"octlb_last_error"}


def test_device_create_reports_errors_without_crashing():
    # inconsistent tables are rejected before any device work
    topo = abi.TopologyArrays(6, 1, np.zeros(1, np.int32), np.array([[0, 0, 0, 1, 1, 1.0]]),
                              np.array([[0, 0, 5]], np.int32), [], np.zeros((0, 2), np.int32))
    p = abi.make_params(1.5)
    import ctypes as C
    h = C.c_void_p()
    rc = abi.lib().octlb_dev_create(C.byref(topo.struct()), C.byref(p), C.byref(h))
    assert rc != 0 and not h.value
    assert abi.lib().octlb_last_error()


def test_adjoint_and_output_entry_points_reject_null_arguments():
    # argument checks run before any device work
    import ctypes as C
    lib = abi.lib()
    h = C.c_void_p()
    assert lib.octlb_adj_create(None, C.byref(h)) == -1 and not h.value
    z = np.zeros(6)
    rc = lib.octlb_adj_run(None, 10, 0, 10, 5, abi.f64p(z), abi.f64p(z), 1.0, abi.f64p(z), 1, 0, None,
                           abi.f64p(z), abi.f64p(z), None)
    assert rc == -1 and b"null" in lib.octlb_last_error()
    assert lib.octlb_adj_destroy(None) == 0
    assert lib.octlb_dev_macro_fields(None, None, None) != 0
    assert lib.octlb_dev_write_vtk(None, 0, b".", b"x") != 0
    assert lib.octlb_vtk_append_pvd(None, 0, b"x.vtm") == -1


def _leaves(topo):
    return topo.bounds


def _brute_faces(topo):
    """Same-level / coarse-fine / boundary faces from geometry only."""
    b = topo.bounds
    lv = topo.level
    dom_lo = b[:, :3].min(0)
    dom_hi = b[:, 3:].max(0)
    same, cf, tb = set(), {}, set()
    tol = 1e-12
    for i in range(len(b)):
        for d in range(6):
            ax, side = d // 2, d & 1
            face = b[i, 3 + ax] if side else b[i, ax]
            if abs(face - (dom_hi[ax] if side else dom_lo[ax])) < tol:
                tb.add((i, d))
                continue
            t = [a for a in range(3) if a != ax]
            for j in range(len(b)):
                if j == i:
                    continue
                other = b[j, ax] if side else b[j, 3 + ax]
                if abs(other - face) > tol:
                    continue
                ov = all(b[j, a] < b[i, 3 + a] - tol and b[j, 3 + a] > b[i, a] + tol for a in t)
                if not ov:
                    continue
                if lv[j] == lv[i]:
                    same.add((i, d, j))
                elif lv[j] == lv[i] + 1:
                    cf.setdefault((i, d), set()).add(j)
    return same, cf, tb


def _forest(kind):
    if kind == "uniform":
        f = Forest([0, 0, 0, 1, 1, 1])
        f.refine_uniform(2)
    elif kind == "center":
        f = Forest([0, 0, 0, 1, 1, 1])
        f.refine_uniform(1)
        f.refine_box([0.4, 0.4, 0.4, 0.6, 0.6, 0.6], 3)
        f.balance()
    else:
        f = Forest([0, 0, 0, 2, 1, 1], (2, 1, 1))
        f.refine_uniform(1)
        f.refine_box([0.3, 0.2, 0.0, 0.9, 0.6, 1.0], 2)
        f.balance()
    return f


def test_uniform_level2_has_64_tiling_octants():
    f = _forest("uniform")
    topo = f.topology(8, 1)
    assert topo.num_blocks == 64 and set(topo.level) == {2}
    vol = np.prod(topo.bounds[:, 3:] - topo.bounds[:, :3], axis=1)
    assert abs(vol.sum() - 1.0) < 1e-12
    assert topo.num_cf == 0
    # uniform forest: only same-level faces; each interior face twice
    assert topo.same_level.shape[0] == 2 * 3 * 4 * 4 * 3


@pytest.mark.parametrize("kind", ["uniform", "center", "brick"])
def test_face_lists_match_brute_force(kind):
    topo = _forest(kind).topology(8, 1)
    same, cf, tb = _brute_faces(topo)
    assert {tuple(r) for r in topo.same_level} == same
    assert {tuple(r) for r in topo.tree_boundary} == tb
    got = {}
    for k in range(topo.num_cf):
        c = topo.coarse_fine[k]
        got[(c.coarse_id, c.normal)] = set(c.fine_ids)
        assert c.fine_level == c.coarse_level + 1
        assert list(c.coarse_bounds) == list(topo.bounds[c.coarse_id])
    assert got == cf
    assert all(len(v) == 4 for v in got.values())


@pytest.mark.parametrize("kind", ["center", "brick"])
def test_two_to_one_face_balance(kind):
    topo = _forest(kind).topology(8, 1)
    same, cf, tb = _brute_faces(topo)
    # every non-boundary face is same-level, coarse->fine, or the fine side of
    # a recorded coarse-fine face; nothing differs by more than one level
    b, lv = topo.bounds, topo.level
    for i in range(len(b)):
        for j in range(len(b)):
            if i == j:
                continue
            touch = all(b[i, a] <= b[j, 3 + a] + 1e-12 and b[j, a] <= b[i, 3 + a] + 1e-12 for a in range(3))
            shared = sum(abs(min(b[i, 3 + a], b[j, 3 + a]) - max(b[i, a], b[j, a])) > 1e-12 for a in range(3))
            if touch and shared == 2:
                assert abs(int(lv[i]) - int(lv[j])) <= 1


def test_morton_order_within_tree():
    f = Forest([0, 0, 0, 1, 1, 1])
    f.refine_uniform(1)
    topo = f.topology(8, 1)
    lo = topo.bounds[:, :3]
    child = [int(round(2 * x)) + 2 * int(round(2 * y)) + 4 * int(round(2 * z)) for x, y, z in lo]
    assert child == list(range(8))  # p8est child id = x + 2y + 4z


def test_coarse_fine_faces_sorted_by_coarse_level():
    f = Forest([0, 0, 0, 1, 1, 1])
    f.refine_uniform(1)
    f.refine_box([0.1, 0.1, 0.1, 0.2, 0.2, 0.2], 3)
    f.balance()
    topo = f.topology(8, 1)
    levels = [topo.coarse_fine[k].coarse_level for k in range(topo.num_cf)]
    assert levels == sorted(levels) and len(set(levels)) >= 2


@pytest.mark.parametrize("field,value,needle", [
    ("normal", 7, b"normal"),
    ("normal", -1, b"normal"),
    ("coarse_level", 0, b"levels"),
    ("fine_level", 3, b"levels"),
    ("coarse_level", 99, b"levels"),
])
def test_create_rejects_inconsistent_coarse_fine_faces(field, value, needle):
    # a face whose normal or levels disagree with the block table is refused
    # before any device work (no out-of-range plan rows, no wrong parities)
    import ctypes as C
    from octlb_amd import cases

    case = cases.cylinder_channel(bricks_x=2, base_level=1, fine=True, xc=0.5, radius=0.1,
                                  refine_lo=0.25, refine_hi=1.0)
    topo = case.topo
    assert topo.num_cf > 0
    setattr(topo.coarse_fine[0], field, value)
    h = C.c_void_p()
    rc = abi.lib().octlb_dev_create(C.byref(topo.struct()), C.byref(case.params(64)), C.byref(h))
    assert rc != 0 and not h.value
    assert needle in abi.lib().octlb_last_error()
