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
