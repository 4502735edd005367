"""Drag-coefficient log in the reference cylinder example's format.

cylinder3d.cpp:104-121 opens ``<output_dir>/cd.csv`` with the header
``iT,Cd`` and appends ``step,Cd`` every cd_interval steps, the double written
by a default C++ stream (six significant digits, %g).  Cd =
2 (-F . flow_dir) / (rho0 u_inlet^2 A) (Cylinder3dCase::drag_coefficient,
cylinder3d_case.h:510-517) from the device drag (octlb_dev_drag).  VTK output
is GpuSolver.write_vtk (octlb_dev_write_vtk).
"""
from __future__ import annotations

from pathlib import Path

import numpy as np


def drag_coefficient(force, rho0: float, u_ref: float, area: float, flow_dir=(1.0, 0.0, 0.0)) -> float:
    f = np.asarray(force, dtype=np.float64)
    d = np.asarray(flow_dir, dtype=np.float64)
    fdot = f[0] * d[0] + f[1] * d[1] + f[2] * d[2]
    return 2.0 * (-fdot) / (rho0 * u_ref * u_ref * area)


class CdCsv:
    def __init__(self, output_dir):
        self.path = Path(output_dir) / "cd.csv"
        self.path.write_text("iT,Cd\n")

    def append(self, step: int, cd: float) -> None:
        with self.path.open("a") as f:
            f.write(f"{int(step)},{cd:g}\n")


def write_amr_timestep(sim, it: int, output_dir, base_name: str = "cavity3d", phys_scale: float = 1.0):
    """AmrVtkWriter::WriteTimestep + WriteVtmAndPvd (amr_vtk_writer.h:47-66,
    amr_vtk_writer.cpp:32-48) as the cavity case uses it
    (cavity3d_case.h:374-387): per block ``<base>_r0_oct<id>_T<it:05d>.vts``
    with CellData "velocity" and "physVelocity" (= velocity * phys_scale,
    the converter's char_phys_velocity / char_lattice_velocity), the
    ``<base>_T<it:05d>.vtm`` index and a DataSet appended to ``<base>.pvd``.
    Each block writes its own fields (the reference's writer binds one
    lattice's fields for every octant, which only matters for several
    blocks).  Returns the .vts paths."""
    import ctypes as C

    from . import abi

    lib = abi.lib()
    out = Path(output_dir)
    out.mkdir(parents=True, exist_ok=True)
    _, u = sim.macro_fields()
    bounds = np.ascontiguousarray(sim.topo.bounds, dtype=np.float64).reshape(-1, 6)
    tag = f"{int(it):05d}"
    names = (C.c_char_p * 2)(b"velocity", b"physVelocity")
    comps = np.array([3, 3], np.int32)
    paths, rel = [], []
    for b in range(sim.nb):
        vel = np.ascontiguousarray(u[b].reshape(-1))
        phys = np.ascontiguousarray(vel * float(phys_scale))
        ptrs = (abi._F64P * 2)(abi.f64p(vel), abi.f64p(phys))
        name = f"{base_name}_r0_oct{b}_T{tag}.vts"
        bb = np.ascontiguousarray(bounds[b])
        abi.check(lib.octlb_vtk_write_vts(str(out / name).encode(), sim.n, sim.n, sim.n, abi.f64p(bb), 2, names,
                                          abi.i32p(comps), ptrs), "write_vts")
        paths.append(out / name)
        rel.append(name.encode())
    arr = (C.c_char_p * max(len(rel), 1))(*rel)
    vtm = f"{base_name}_T{tag}.vtm"
    abi.check(lib.octlb_vtk_write_vtm(str(out / vtm).encode(), len(rel), arr), "write_vtm")
    abi.check(lib.octlb_vtk_append_pvd(str(out / f"{base_name}.pvd").encode(), int(it), vtm.encode()), "append_pvd")
    return paths


# Ghia, Ghia & Shin (1982), Re = 100: u/u_lid on the vertical centreline at
# these y/H (the published table the reference writes beside its samples,
# cavity3d_case.h:108-125)
GHIA_RE100_Y_OVER_H = (1.0, 0.9766, 0.9688, 0.9609, 0.9531, 0.8516, 0.7344, 0.6172, 0.5, 0.4531, 0.2813, 0.1719,
                       0.1016, 0.0703, 0.0625, 0.0547, 0.0)
GHIA_RE100_U_OVER_LID = (1.0, 0.84123, 0.78871, 0.73722, 0.68717, 0.23151, 0.00332, -0.13641, -0.20581, -0.21090,
                         -0.15662, -0.10150, -0.06434, -0.04775, -0.04192, -0.03717, 0.0)


def cavity_ux_field(sim, case):
    """Cavity3dCase::lattice_ux_at (cavity3d_case.h:205-218) for every cell
    of the single-block cavity, [ix][iy][iz]: the prescribed wall velocity on
    velocity-Dirichlet cells (the lid on interior top-face cells only,
    IsInteriorTopLidCell, interpolated_velocity.h:64-117), CellProxy u_x
    elsewhere (from the device)."""
    from . import abi

    if sim.nb != 1:
        raise ValueError("the centreline sampler is the single-block cavity's")
    n = sim.n
    _, u = sim.macro_fields()
    ux = np.array(u[0][..., 0].transpose(2, 1, 0))  # [k][j][i] -> [i][j][k]
    lid = 0.0
    for sp in case.specs:
        if sp.face == abi.YMAX:
            lid = float(sp.u_wall[0])
            break
    wall = np.zeros_like(ux)
    wall[1:n - 1, n - 1, 1:n - 1] = lid
    return np.where(case.kinds[0] == abi.KIND_VELOCITY_DIRICHLET, wall, ux)


def cavity_centerline(ux, resolution: int, u_lid: float, length: float = 1.0):
    """Cavity3dCase::sample_ghia_centerline (cavity3d_case.h:325-338): u_x/u_lid
    at (L/2, y/H L, L/2) by the trilinear rule of sample_ux (:242-286),
    lattice index floor(coord/dx) clamped to [0, n-2], dx = L/resolution."""
    n = ux.shape[0]
    dx = length / float(resolution)

    def lo(c):
        return min(max(int(np.floor(c / dx)), 0), n - 2)

    out = []
    x = z = 0.5 * length
    for yh in GHIA_RE100_Y_OVER_H:
        y = yh * length
        i0, j0, k0 = lo(x), lo(y), lo(z)
        wx, wy, wz = (x - i0 * dx) / dx, (y - j0 * dx) / dx, (z - k0 * dx) / dx
        c = ux[i0:i0 + 2, j0:j0 + 2, k0:k0 + 2]
        c00 = c[0, 0, 0] * (1.0 - wx) + c[1, 0, 0] * wx
        c10 = c[0, 1, 0] * (1.0 - wx) + c[1, 1, 0] * wx
        c01 = c[0, 0, 1] * (1.0 - wx) + c[1, 0, 1] * wx
        c11 = c[0, 1, 1] * (1.0 - wx) + c[1, 1, 1] * wx
        c0 = c00 * (1.0 - wy) + c10 * wy
        c1 = c01 * (1.0 - wy) + c11 * wy
        out.append(float(c0 * (1.0 - wz) + c1 * wz) / u_lid)
    return out


def write_centerline_csv(path, simulated) -> None:
    """Cavity3dCase::write_centerline_csv (cavity3d_case.h:389-397): header
    y_over_H,u_over_lid,ghia, doubles as a default C++ stream writes them."""
    with Path(path).open("w") as f:
        f.write("y_over_H,u_over_lid,ghia\n")
        for y, s, g in zip(GHIA_RE100_Y_OVER_H, simulated, GHIA_RE100_U_OVER_LID):
            f.write(f"{y:g},{s:g},{g:g}\n")
