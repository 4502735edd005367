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
