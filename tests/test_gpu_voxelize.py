"""octlb_voxelize_part (the device voxelize_part, part_voxelizer.cpp:71-114)
against the oracle's exact statement: every MaterialKind byte equal, on the
degenerate small cases the oracle is pinned on (faces on cell faces, rays
along faces and through edges, collinear triangles), on the reference's
engine tests, and on the benchmarked cylinder forest with its 8,192-triangle
soup; then the whole geometry pipeline on the device, soup -> material ->
Bouzidi links, against the oracle's pipeline."""
import numpy as np
import pytest

from octlb_amd import abi, cases
from octlb_amd import geometry as geo
from oracle import binding

from test_oracle_voxelize import CASES, CH, OB, _forest, _uniform, material_at, tunnel

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", sorted(CASES))
def test_device_matches_oracle_small(name):
    role, soup, inner = CASES[name]
    topo = _forest()
    print(f"voxelize {name}", flush=True)
    got, nexact = geo.voxelize_part(topo.bounds, topo.n, soup, role, inner)
    want = binding.voxelize_part(topo.bounds, topo.n, soup, role, inner)
    np.testing.assert_array_equal(got, want)
    print(f"  {nexact} cells took the exact fallback", flush=True)


def test_exact_fallback_is_exercised():
    # faces on cell boundaries leave the filters undecided on touching cells
    role, soup, inner = CASES["fixture_box_on_faces"]
    topo = _forest()
    _, nexact = geo.voxelize_part(topo.bounds, topo.n, soup, role, inner)
    assert nexact > 0


def test_reference_engine_assertions_on_device():
    topo = _uniform()
    cube = geo.reference_box_soup((0.4,) * 3, (0.6,) * 3)
    s1 = geo.build_material(topo.bounds, topo.n, [dict(soup=cube, role=OB, priority=0)])
    assert (s1 == geo.MAT_SOLID).any() and (s1 == geo.MAT_BOUNDARY).any()
    assert material_at(topo, s1, 0.05, 0.05, 0.05) == geo.MAT_FLUID
    assert material_at(topo, s1, 0.5, 0.5, 0.5) in (geo.MAT_SOLID, geo.MAT_BOUNDARY)
    hi = geo.build_material(topo.bounds, topo.n, tunnel(10))
    assert material_at(topo, hi, 0.5, 0.5, 0.5) == geo.MAT_SOLID
    low = tunnel(0)
    low[0]["priority"], low[1]["priority"] = 10, 0
    assert material_at(topo, geo.build_material(topo.bounds, topo.n, low), 0.5, 0.5, 0.5) == geo.MAT_FLUID
    want = geo.build_material(topo.bounds, topo.n, tunnel(10),
                              voxelize=lambda b, m, s, r, i: binding.voxelize_part(b, m, s, r, i))
    np.testing.assert_array_equal(hi, want)
    with pytest.raises(abi.OctlbError, match="requires inner_cavity_soup"):
        geo.voxelize_part(topo.bounds, topo.n, cube, CH)
    with pytest.raises(abi.OctlbError, match="no triangles"):
        geo.voxelize_part(topo.bounds, topo.n, np.zeros((0, 9)), OB)


def _cylinder():
    case = cases.cylinder_channel(bricks_x=16, bricks_yz=4, base_level=0, fine=True, xc=1.0,
                                  radius=0.125, u_in=0.04, tau=0.55)
    return case.topo, geo.cylinder_soup(1.0, 0.5, 0.125, 0.0, 1.0, 512, 8)


def test_cylinder_forest_matches_oracle_and_geometry():
    topo, soup = _cylinder()
    print(f"voxelize cylinder: {topo.num_blocks} blocks, {soup.shape[0]} triangles", flush=True)
    got, nexact = geo.voxelize_part(topo.bounds, 8, soup, OB)
    want = binding.voxelize_part(topo.bounds, 8, soup, OB)
    np.testing.assert_array_equal(got, want)
    # the polygon lies between the circles of radius r cos(pi/512) and r:
    # cells whose box clears that band are solid inside, fluid outside
    b = topo.bounds
    idx = (np.arange(8) + 0.5) / 8
    w = b[:, 3:] - b[:, :3]
    X = b[:, 0, None, None, None] + w[:, 0, None, None, None] * idx[None, None, None, :]
    Y = b[:, 1, None, None, None] + w[:, 1, None, None, None] * idx[None, None, :, None]
    d = np.sqrt((X - 1.0) ** 2 + (Y - 0.5) ** 2) + 0 * idx[None, :, None, None]
    half = (w[:, 0] / 8 * np.sqrt(3.0) / 2)[:, None, None, None] + 0 * d
    r = 0.125
    assert (got[d + half < r * np.cos(np.pi / 512)] == geo.MAT_SOLID).all()
    assert (got[d - half > r] == geo.MAT_FLUID).all()
    assert (got == geo.MAT_BOUNDARY).sum() > 0
    print(f"  {nexact} cells took the exact fallback", flush=True)


def test_soup_to_links_pipeline_matches_oracle():
    # GeometryEngine::build's material, then BouzidiLinkData::Build, both on
    # the device, against the oracle's pipeline: identical links and q
    topo, soup = _cylinder()
    mat, _ = geo.voxelize_part(topo.bounds, 8, soup, OB)
    link, q = geo.build_links(topo.bounds, 8, mat, soup)
    omat = binding.voxelize_part(topo.bounds, 8, soup, OB)
    np.testing.assert_array_equal(mat, omat)
    # only blocks holding solid-like cells can have links: the oracle runs
    # on those (its every-cell-every-triangle loop), block ids mapped back
    sel = np.flatnonzero((omat != geo.MAT_FLUID).reshape(topo.num_blocks, -1).any(1))
    olink, oq = binding.build_links(topo.bounds[sel], 8, omat[sel].reshape(-1), soup)
    olink = olink.copy()
    olink[:, 0] = sel[olink[:, 0]]
    np.testing.assert_array_equal(link, olink)
    np.testing.assert_array_equal(q, oq)
    assert len(q) > 1000
