"""Device Bouzidi link build (octlb_build_links) against the oracle's
restatement of BouzidiLinkData::Build (bouzidi_link_data.cpp:139-233):
links and q must be bit-identical -- the same cells, populations and FP64
fractions, in (block, cell, population) order.  Cases: the reference's
single-triangle wall, a sphere over a 2-level forest (cell widths differ per
block), a z-spanning cylinder over the cylinder-channel forest with more
triangles than one LDS chunk, duplicated triangles (ties), the empty soup,
and the count query / capacity errors of the C ABI; and the whole pipeline:
device-built links driving the 2-level cylinder step, against the oracle.
"""
import ctypes as C

import numpy as np
import pytest

from octlb_amd import abi, cases, geometry
from octlb_amd.forest import Forest
from oracle.binding import build_links as oracle_links

pytestmark = pytest.mark.gpu


def _same(bounds, n, mat, tris):
    gl, gq = geometry.build_links(bounds, n, mat, tris)
    ol, oq = oracle_links(bounds, n, mat, tris)
    print(f"links: {len(gl)}", flush=True)
    np.testing.assert_array_equal(gl, ol)
    np.testing.assert_array_equal(gq, oq)
    return gl, gq


def test_reference_wall_case():
    mat = np.zeros((1, 8, 8, 8), np.uint8)
    mat[..., 4:] = geometry.MAT_SOLID
    wall = np.array([[0.5, 0.0, 0.0, 0.5, 1.0, 0.0, 0.5, 0.0, 1.0]])
    link, q = _same(np.array([[0, 0, 0, 1, 1, 1.0]]), 8, mat, wall)
    assert len(link) > 0 and np.all(link[:, 1] // 64 == 3)


def test_sphere_two_level_forest_with_ties():
    f = Forest([0, 0, 0, 1, 1, 1])
    f.refine_uniform(1)
    f.refine_box([0.25, 0.25, 0.25, 0.75, 0.75, 0.75], 2)
    f.balance()
    topo = f.topology(8, 1)
    ctr, r = (0.5, 0.47, 0.52), 0.21
    mat = geometry.material_field(topo, lambda x, y, z: (x - ctr[0]) ** 2 + (y - ctr[1]) ** 2 + (z - ctr[2]) ** 2 < r * r)
    tris = geometry.sphere_soup(ctr, r, 20)
    tris = np.concatenate([tris[:50], tris, tris[:50]])
    link, _ = _same(topo.bounds, 8, mat, tris)
    assert len(link) > 1000


def test_cylinder_channel_forest():
    case = cases.cylinder_channel(bricks_x=4, base_level=1, fine=True, xc=1.0, radius=0.125)
    topo = case.topo
    mat = geometry.material_field(topo, lambda x, y, z: (x - 1.0) ** 2 + (y - 0.5) ** 2 < 0.125 ** 2)
    tris = geometry.cylinder_soup(1.0, 0.5, 0.125, 0.0, 1.0, 160, rings=3)  # 960 triangles
    link, q = _same(topo.bounds, 8, mat, tris)
    assert len(link) > 1000 and np.all((q > 1e-8) & (q < 1 - 1e-8))


def test_empty_soup_count_query_and_capacity():
    f = Forest([0, 0, 0, 1, 1, 1])
    f.refine_uniform(1)
    topo = f.topology(8, 1)
    mat = geometry.material_field(topo, lambda x, y, z: x > 0.55)
    gl, gq = geometry.build_links(topo.bounds, 8, mat, np.zeros((0, 9)))
    assert len(gl) == 0 and len(gq) == 0
    h = abi.lib()
    tris = np.ascontiguousarray(geometry.box_soup((0.55, -1, -1), (2, 2, 2)))
    b = np.ascontiguousarray(topo.bounds)
    m = np.ascontiguousarray(mat.reshape(-1))
    cnt = C.c_int64(-1)
    args = (topo.num_blocks, 8, abi.f64p(b), abi.u8p(m), tris.shape[0], abi.f64p(tris))
    assert h.octlb_build_links(*args, 0, C.byref(cnt), None, None) == 0
    assert cnt.value > 0
    small = np.zeros((1, 3), np.int32)
    qs = np.zeros(1)
    rc = h.octlb_build_links(*args, 0, C.byref(cnt), abi.i32p(small), abi.f64p(qs))
    assert rc == abi.EINVAL if hasattr(abi, "EINVAL") else rc < 0
    assert b"capacity" in h.octlb_last_error()


def test_device_links_drive_the_solver():
    """Geometry -> q -> step: links built on the device from a triangulated
    cylinder are set on the 2-level cylinder case (kinds from the same
    material field) and the solver matches the oracle run with those links
    to 1e-12, drag included."""
    from dataclasses import replace

    from tests.util import make_gpu, make_oracle, rel_err

    case = cases.cylinder_channel(bricks_x=2, base_level=1, fine=True, xc=0.5, radius=0.1,
                                  refine_lo=0.25, refine_hi=1.0)
    inside = lambda x, y, z: (x - 0.5) ** 2 + (y - 0.5) ** 2 < 0.1 ** 2  # noqa: E731
    mat = geometry.material_field(case.topo, inside)
    kinds = cases.material_to_kinds(mat.transpose(0, 3, 2, 1).astype(np.int64))
    cases.stamp_tree_boundary(kinds, case.topo.tree_boundary, case.specs)
    tris = geometry.cylinder_soup(0.5, 0.5, 0.1, 0.0, 1.0, 96, rings=2)
    link, q = geometry.build_links(case.topo.bounds, 8, mat, tris)
    ol, oq = oracle_links(case.topo.bounds, 8, mat, tris)
    np.testing.assert_array_equal(link, ol)
    np.testing.assert_array_equal(q, oq)
    assert len(link) > 100
    c2 = replace(case, kinds=kinds, link=link, q=q)
    g, o = make_gpu(c2), make_oracle(c2)
    g.advance(10)
    o.advance(10)
    assert rel_err(g.download_core(), o.download_core()) <= 1e-12
    fd, fo = g.drag(c2.drag_bbox), o.drag(c2.drag_bbox)
    assert np.abs(fd - fo).max() <= 1e-12 * max(np.abs(fo).max(), 1e-30)
