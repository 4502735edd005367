"""VTK output in the reference's formats (octlb_vtk_*): the library's
writers against the reference's own (structured_grid_writer.cpp,
vtk_binary_codec.cpp, parallel_vtk_index.cpp compiled into oracle/_ref/ref_vtk
from the reference sources) byte for byte, and the reference's unit cases
(test_vtk_writer.cpp:83-139: constant scalar on 2^3 cells with 27 points,
corner bounds on a 4^3 block, a scalar + vector pair) through an independent
base64 decode.  Host-only calls: no GPU needed.

A plausible slip -- the byte count missing from the binary header, padding
dropped from the base64 tail, x-slowest cell order, corner points from
linspace instead of the reference's a + (b - a) i / n, the .pvd closing tags
not replaced -- fails one of them.
"""
import base64
import ctypes as C
import re
import struct
import subprocess
from pathlib import Path

import numpy as np
import pytest

from octlb_amd import abi

REF_VTK = Path(__file__).resolve().parents[1] / "oracle" / "_ref" / "ref_vtk"
need_ref = pytest.mark.skipif(not REF_VTK.exists(), reason="reference VTK writer not built (no reference sources)")


def _write_vts(path, n3, bounds, fields):
    lib = abi.lib()
    names = (C.c_char_p * max(len(fields), 1))(*[f[0].encode() for f in fields])
    comps = np.array([f[1] for f in fields] or [1], np.int32)
    arrs = [np.ascontiguousarray(f[2], dtype=np.float64).reshape(-1) for f in fields]
    ptrs = (abi._F64P * max(len(fields), 1))(*[abi.f64p(a) for a in arrs])
    b = np.ascontiguousarray(bounds, dtype=np.float64)
    abi.check(lib.octlb_vtk_write_vts(str(path).encode(), n3[0], n3[1], n3[2], abi.f64p(b), len(fields), names,
                                      abi.i32p(comps), ptrs), "write_vts")


def _ref_vts(tmp_path, path, n3, bounds, fields):
    blob = struct.pack("<3i", *n3) + struct.pack("<6d", *bounds) + struct.pack("<i", len(fields))
    for name, comp, data in fields:
        blob += struct.pack("<i", len(name)) + name.encode() + struct.pack("<i", comp)
        blob += np.ascontiguousarray(data, dtype="<f8").tobytes()
    inp = tmp_path / "in.bin"
    inp.write_bytes(blob)
    subprocess.run([str(REF_VTK), "vts", str(inp), str(path)], check=True, timeout=60)


def _decode(xml, name=None):
    if name is None:
        m = re.search(r'<Points>\s*<DataArray[^>]*>\s*([A-Za-z0-9+/=]+)', xml)
    else:
        m = re.search(r'Name="%s"[^>]*>\s*([A-Za-z0-9+/=]+)' % re.escape(name), xml)
    raw = base64.b64decode(m.group(1))
    nbytes = struct.unpack("<I", raw[:4])[0]
    assert nbytes == len(raw) - 4
    return np.frombuffer(raw[4:], dtype="<f8")


@need_ref
@pytest.mark.parametrize("n3", [(8, 8, 8), (3, 5, 2), (1, 1, 1)])
def test_vts_bytes_match_the_reference_writer(tmp_path, n3):
    rng = np.random.default_rng(sum(n3))
    nc = n3[0] * n3[1] * n3[2]
    bounds = [0.125, -0.5, 0.3, 0.625, 0.1, 0.71]
    fields = [("velocity", 3, rng.standard_normal(nc * 3)), ("pressure", 1, rng.random(nc) / 3.0),
              ("two", 2, rng.standard_normal(nc * 2))]
    for k in range(len(fields) + 1):  # 0..3 fields: every base64 padding case occurs
        ours, ref = tmp_path / f"o{k}.vts", tmp_path / f"r{k}.vts"
        _write_vts(ours, n3, bounds, fields[:k])
        _ref_vts(tmp_path, ref, n3, bounds, fields[:k])
        assert ours.read_bytes() == ref.read_bytes(), k


@need_ref
def test_vtm_and_pvd_match_the_reference(tmp_path):
    lib = abi.lib()
    (tmp_path / "o").mkdir()
    (tmp_path / "r").mkdir()
    rel = [f"cyl_block{b}_T40.vts" for b in range(3)]
    arr = (C.c_char_p * 3)(*[r.encode() for r in rel])
    abi.check(lib.octlb_vtk_write_vtm(str(tmp_path / "o" / "cyl_T00040.vtm").encode(), 3, arr), "vtm")
    subprocess.run([str(REF_VTK), "vtm", str(tmp_path / "r"), "cyl", "40"] + rel, check=True, timeout=60)
    assert (tmp_path / "o" / "cyl_T00040.vtm").read_bytes() == (tmp_path / "r" / "cyl_T00040.vtm").read_bytes()
    for it in (40, 80, 120):
        abi.check(lib.octlb_vtk_append_pvd(str(tmp_path / "o" / "cyl.pvd").encode(), it, f"cyl_T{it}.vtm".encode()),
                  "pvd")
        subprocess.run([str(REF_VTK), "pvd", str(tmp_path / "r" / "cyl.pvd"), str(it), f"cyl_T{it}.vtm"], check=True,
                       timeout=60)
        assert (tmp_path / "o" / "cyl.pvd").read_bytes() == (tmp_path / "r" / "cyl.pvd").read_bytes(), it


def test_reference_unit_cases_decode(tmp_path):
    # SingleBlock_ConstantScalar: 2^3 cells of 3.5 over the unit cube
    p = tmp_path / "a.vts"
    _write_vts(p, (2, 2, 2), [0, 0, 0, 1, 1, 1], [("dummy_scalar", 1, np.full(8, 3.5))])
    xml = p.read_text()
    assert 'NumberOfPoints="27"' in xml and 'NumberOfCells="8"' in xml and "StructuredGrid" in xml
    np.testing.assert_array_equal(_decode(xml, "dummy_scalar"), np.full(8, 3.5))
    # SingleBlock_CornerBoundsMatch: first and last points are the box corners
    bounds = [0.2, 0.3, 0.4, 0.8, 0.9, 1.0]
    p = tmp_path / "b.vts"
    _write_vts(p, (4, 4, 4), bounds, [("s", 1, np.zeros(64))])
    pts = _decode(p.read_text()).reshape(-1, 3)
    assert len(pts) == 125
    np.testing.assert_array_equal(pts[0], bounds[:3])
    np.testing.assert_allclose(pts[-1], bounds[3:], rtol=0, atol=1e-15)  # ApproxEq, as the reference test
    # x fastest, then y, then z; a + (b - a) * i / n per axis
    i = np.arange(5)
    xs = bounds[0] + (bounds[3] - bounds[0]) * i / 4
    np.testing.assert_array_equal(pts[:5, 0], xs)
    np.testing.assert_array_equal(pts[:5, 1], np.full(5, bounds[1]))
    # SingleBlock_MultiField: scalar + vector, cell values in VTK order
    vec = np.arange(24, dtype=float)
    p = tmp_path / "c.vts"
    _write_vts(p, (2, 2, 2), [0, 0, 0, 1, 1, 1], [("dummy_scalar", 1, np.ones(8)), ("dummy_vector", 3, vec)])
    xml = p.read_text()
    assert 'Name="dummy_scalar"' in xml and 'Name="dummy_vector"' in xml and 'NumberOfComponents="3"' in xml
    np.testing.assert_array_equal(_decode(xml, "dummy_vector"), vec)


def test_writer_errors_are_reported(tmp_path):
    lib = abi.lib()
    b = np.zeros(6)
    rc = lib.octlb_vtk_write_vts(str(tmp_path / "x.vts").encode(), 0, 1, 1, abi.f64p(b), 0, None, None, None)
    assert rc == -1  # OCTLB_EINVAL: non-positive extent
    rc = lib.octlb_vtk_write_vtm(str(tmp_path / "missing" / "x.vtm").encode(), 0, None)
    assert rc == -5 and b"failed to open" in lib.octlb_last_error()  # OCTLB_EIO


def test_cd_csv_format(tmp_path):
    from octlb_amd.output import CdCsv, drag_coefficient

    log = CdCsv(tmp_path)
    # default C++ stream output of these doubles (precision 6, %g)
    for step, v in ((100, 6.3612345), (200, 1e-7), (300, 123456789.0), (400, -0.5)):
        log.append(step, v)
    assert log.path.read_text() == "iT,Cd\n100,6.36123\n200,1e-07\n300,1.23457e+08\n400,-0.5\n"
    # Cylinder3dCase::drag_coefficient: 2 (-F . dir) / (rho0 u^2 A)
    assert drag_coefficient((-0.5, 0.1, 0.0), 1.0, 0.05, 64.0) == 2.0 * 0.5 / (0.05 * 0.05 * 64.0)


def test_centerline_sampler_and_csv(tmp_path):
    from octlb_amd.output import GHIA_RE100_Y_OVER_H, cavity_centerline, write_centerline_csv

    # trilinear sampling reproduces a linear field exactly (away from clamping)
    n, res = 31, 30
    i = np.arange(n, dtype=float)
    ux = (0.5 + 0.25 * i[:, None, None] + 2.0 * i[None, :, None] - 0.75 * i[None, None, :]) * np.ones((n, n, n))
    prof = cavity_centerline(ux, res, u_lid=2.0)
    dx = 1.0 / res
    for y, p in zip(GHIA_RE100_Y_OVER_H, prof):
        want = 0.5 + 0.25 * (0.5 / dx) + 2.0 * min(y / dx, n - 1) - 0.75 * (0.5 / dx)
        assert p == pytest.approx(want / 2.0, rel=1e-12, abs=1e-12)
    write_centerline_csv(tmp_path / "c.csv", prof)
    lines = (tmp_path / "c.csv").read_text().splitlines()
    assert lines[0] == "y_over_H,u_over_lid,ghia" and len(lines) == 18
    assert lines[1].startswith("1,") and lines[1].endswith(",1") and lines[-1].startswith("0,")
