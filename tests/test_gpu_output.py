"""Output from device buffers: octlb_dev_macro_fields against CellProxy's
density and velocity (block_lattice.h:113-125) recomputed from the
downloaded populations in the same operation order (bit for bit in FP64),
and octlb_dev_write_vtk's files (the cylinder case's per-block velocity and
pressure .vts, the .vtm index, the .pvd collection) decoded back to those
fields.
"""
import base64
import re
import struct

import numpy as np
import pytest

from octlb_amd import cases
from octlb_amd.solver import GpuSolver
from oracle.binding import setup

pytestmark = pytest.mark.gpu

C = cases.C


def _proxy_rho_u(core):
    """CellProxy::computeRhoU: rho starts at 1, populations added in order."""
    f = core.reshape(-1, 19)
    rho = np.ones(len(f))
    u = np.zeros((len(f), 3))
    for i in range(19):
        rho = rho + f[:, i]
        for d in range(3):
            if C[i, d] != 0:
                u[:, d] = u[:, d] + f[:, i] * float(C[i, d])
    return rho, u / rho[:, None]


@pytest.fixture(scope="module")
def stepped():
    case = cases.cylinder_channel(bricks_x=2, base_level=1, fine=True, u_in=0.05)
    sim = GpuSolver(case.topo, case.params(64))
    setup(sim, case)
    sim.advance(6)
    return case, sim


def test_macro_fields_are_cellproxy_rho_u(stepped):
    case, sim = stepped
    print(f"blocks {sim.nb}", flush=True)
    rho, u = sim.macro_fields()
    core = sim.download_core()  # [b][ix][iy][iz][19]
    r, v = _proxy_rho_u(core)
    n = sim.n
    r = r.reshape(sim.nb, n, n, n).transpose(0, 3, 2, 1)  # -> [b][k][j][i]
    v = v.reshape(sim.nb, n, n, n, 3).transpose(0, 3, 2, 1, 4)
    np.testing.assert_array_equal(rho, r)
    np.testing.assert_array_equal(u, v)
    assert np.abs(u).max() > 0


def _decode(xml, name):
    m = re.search(r'Name="%s"[^>]*>\s*([A-Za-z0-9+/=]+)' % name, xml)
    raw = base64.b64decode(m.group(1))
    assert struct.unpack("<I", raw[:4])[0] == len(raw) - 4
    return np.frombuffer(raw[4:], dtype="<f8")


def test_write_vtk_files(stepped, tmp_path):
    case, sim = stepped
    rho, u = sim.macro_fields()
    sim.write_vtk(6, tmp_path, "cyl")
    sim.write_vtk(12, tmp_path, "cyl")
    for b in (0, sim.nb - 1):
        xml = (tmp_path / f"cyl_block{b}_T6.vts").read_text()
        np.testing.assert_array_equal(_decode(xml, "velocity"), u[b].reshape(-1))
        np.testing.assert_array_equal(_decode(xml, "pressure"), rho[b].reshape(-1) * (1.0 / 3.0))
    vtm = (tmp_path / "cyl_T00006.vtm").read_text()
    assert vtm.count("<Block index=") == sim.nb and 'file="cyl_block0_T6.vts"' in vtm
    pvd = (tmp_path / "cyl.pvd").read_text()
    assert pvd.count("<DataSet timestep=") == 2 and 'file="cyl_T12.vtm"' in pvd
    assert pvd.endswith("</Collection>\n</VTKFile>\n")


def test_fp32_macro_fields(stepped):
    case, _ = stepped
    sim = GpuSolver(case.topo, case.params(32))
    setup(sim, case)
    sim.advance(6)
    rho, u = sim.macro_fields()
    r, v = _proxy_rho_u(sim.download_core())
    n = sim.n
    np.testing.assert_allclose(rho, r.reshape(sim.nb, n, n, n).transpose(0, 3, 2, 1), rtol=1e-6)
    np.testing.assert_allclose(u, v.reshape(sim.nb, n, n, n, 3).transpose(0, 3, 2, 1, 4), rtol=0, atol=1e-6)
    sim.close()


def test_amr_writer_cavity_style(tmp_path):
    from octlb_amd.output import write_amr_timestep

    case = cases.cavity_blocked(16, tau=0.6, u_lid=0.05, const_rho=0)
    sim = GpuSolver(case.topo, case.params(64))
    setup(sim, case)
    sim.advance(5)
    _, u = sim.macro_fields()
    paths = write_amr_timestep(sim, 5, tmp_path, "cav", phys_scale=2.5)
    assert [p.name for p in paths[:2]] == ["cav_r0_oct0_T00005.vts", "cav_r0_oct1_T00005.vts"]
    xml = paths[1].read_text()
    np.testing.assert_array_equal(_decode(xml, "velocity"), u[1].reshape(-1))
    np.testing.assert_array_equal(_decode(xml, "physVelocity"), u[1].reshape(-1) * 2.5)
    assert (tmp_path / "cav_T00005.vtm").read_text().count("<Block index=") == sim.nb
    assert 'file="cav_T00005.vtm"' in (tmp_path / "cav.pvd").read_text()
    sim.close()
