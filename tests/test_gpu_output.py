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


# OpenLB cavity3d at its defaults, converged (the profile the reference's
# test_cavity3d_serial.cpp:56-81 compares with, cavity3d_case.h:114-120)
OPENLB_CAVITY3D_U_OVER_LID = (1.0, 0.629597, 0.506129, 0.429369, 0.371218, 0.110531, 0.0733846, 0.0481395,
                              0.00778279, -0.0115753, -0.152547, -0.213294, -0.167816, -0.124812, -0.112694,
                              -0.0995179, 0.0)


def test_cavity_centerline_matches_openlb(tmp_path):
    """Cavity3dSerial.OpenLbCenterline_RelativeL2_Below2Pct: the
    OpenLbCavity3dDefaults cavity (resolution 30, tau 0.509, nu 0.001) after
    5,269 steps, centreline within 2 % (relative L2) of OpenLB's."""
    from octlb_amd.output import cavity_centerline, cavity_ux_field, write_centerline_csv

    from oracle.binding import OracleSim

    case = cases.cavity(n_lat=31, tau=0.509, nu=0.001)
    sim = GpuSolver(case.topo, case.params(64))
    setup(sim, case)
    # an odd block edge: the first steps bit for bit against the oracle
    ora = setup(OracleSim(case.topo, case.params(64)), case)
    sim.advance(3)
    ora.advance(3)
    np.testing.assert_array_equal(sim.download_core(), ora.download_core())
    ora.close()
    print("cavity 31^3, 5269 steps", flush=True)
    sim.advance_graph(5266)
    prof = cavity_centerline(cavity_ux_field(sim, case), case.info["resolution"], case.u_ref)
    sim.close()
    ref = np.array(OPENLB_CAVITY3D_U_OVER_LID)
    l2 = float(np.sqrt(np.sum((np.array(prof) - ref) ** 2) / np.sum(ref ** 2)))
    print(f"centreline relative L2 vs OpenLB: {l2:.4f}", flush=True)
    assert l2 < 0.02
    write_centerline_csv(tmp_path / "centerline.csv", prof)
    assert (tmp_path / "centerline.csv").read_text().count("\n") == 18
