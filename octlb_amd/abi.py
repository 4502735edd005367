"""ctypes mirror of include/octlb_gpu.h and the loader of liboctlb_gpu.so.

The shared library is built in-tree by ``octlb_amd.build`` (hipcc, gfx950).
Importing this module never falls back to anything: if the library is
missing the first call raises ``OctlbError`` with the build command.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

Q = 19
PKG_DIR = Path(__file__).resolve().parent
LIB_PATH = PKG_DIR / "liboctlb_gpu.so"

# FaceDir (src/common/types.h:15-22)
XMIN, XMAX, YMIN, YMAX, ZMIN, ZMAX = range(6)
# BcKind (src/solver/lbm/bc_kind.h:14-23)
(KIND_BULK, KIND_SOLID, KIND_BOUNCE_BACK, KIND_MOVING_BOUNCE_BACK, KIND_BOUZIDI,
 KIND_VELOCITY_DIRICHLET, KIND_PRESSURE_DIRICHLET, KIND_OUTFLOW) = range(8)
# DomainBcType (src/solver/lbm/domain_boundary_handler.h:16-22)
BC_NO_SLIP, BC_MOVING_LID, BC_INTERP_VELOCITY, BC_OUTFLOW, BC_INTERP_PRESSURE = range(5)
INLET_NONE, INLET_UNIFORM, INLET_POISEUILLE, INLET_CHANNEL = range(4)
HANDLER_NOOP, HANDLER_CONCRETE = 0, 1


class OctlbError(RuntimeError):
    pass


class BcSpec(C.Structure):
    _fields_ = [
        ("face", C.c_int32), ("type", C.c_int32), ("u_wall", C.c_double * 3),
        ("rho_target", C.c_double), ("inlet", C.c_int32), ("inlet_face", C.c_int32),
        ("inlet_u", C.c_double * 3), ("u_peak", C.c_double),
        ("n_inplane0", C.c_int32), ("n_inplane1", C.c_int32),
        ("lo0", C.c_double), ("hi0", C.c_double), ("lo1", C.c_double), ("hi1", C.c_double),
        ("ramp_end_t", C.c_double),
    ]


class CfFace(C.Structure):
    _fields_ = [
        ("coarse_id", C.c_int32), ("fine_ids", C.c_int32 * 4), ("normal", C.c_int32),
        ("coarse_level", C.c_int32), ("fine_level", C.c_int32),
        ("coarse_bounds", C.c_double * 6), ("fine_bounds", (C.c_double * 6) * 4),
    ]


class Topology(C.Structure):
    _fields_ = [
        ("num_blocks", C.c_int32), ("n", C.c_int32), ("halo", C.c_int32),
        ("level", C.POINTER(C.c_int32)), ("bounds", C.POINTER(C.c_double)),
        ("num_same_level", C.c_int32), ("same_level", C.POINTER(C.c_int32)),
        ("num_coarse_fine", C.c_int32), ("coarse_fine", C.POINTER(CfFace)),
        ("num_tree_boundary", C.c_int32), ("tree_boundary", C.POINTER(C.c_int32)),
    ]


class Params(C.Structure):
    _fields_ = [
        ("omega", C.c_double), ("precision", C.c_int32), ("handler", C.c_int32),
        ("const_rho", C.c_int32), ("stats_scope", C.c_int32), ("bouzidi", C.c_int32),
        ("snapshot", C.c_int32), ("num_specs", C.c_int32), ("specs", C.POINTER(BcSpec)),
        ("fast_stats", C.c_int32),
    ]


def make_spec(face, type_, u_wall=(0.0, 0.0, 0.0), rho_target=1.0, inlet=INLET_NONE,
              inlet_face=None, inlet_u=(0.0, 0.0, 0.0), u_peak=0.0, n_inplane=(1, 1),
              lo0=0.0, hi0=1.0, lo1=0.0, hi1=1.0, ramp_end_t=0.0) -> BcSpec:
    s = BcSpec()
    s.face = face
    s.type = type_
    s.u_wall[:] = list(u_wall)
    s.rho_target = rho_target
    s.inlet = inlet
    s.inlet_face = face if inlet_face is None else inlet_face
    s.inlet_u[:] = list(inlet_u)
    s.u_peak = u_peak
    s.n_inplane0, s.n_inplane1 = n_inplane
    s.lo0, s.hi0, s.lo1, s.hi1 = lo0, hi0, lo1, hi1
    s.ramp_end_t = ramp_end_t
    return s


class TopologyArrays:
    """Owns numpy copies of a flattened forest and exposes a ``Topology``."""

    def __init__(self, n, halo, level, bounds, same_level, coarse_fine, tree_boundary):
        self.n = int(n)
        self.halo = int(halo)
        self.level = np.ascontiguousarray(level, dtype=np.int32)
        self.bounds = np.ascontiguousarray(bounds, dtype=np.float64).reshape(-1, 6)
        self.same_level = np.ascontiguousarray(same_level, dtype=np.int32).reshape(-1, 3)
        self.coarse_fine = (CfFace * max(1, len(coarse_fine)))()
        for i, cf in enumerate(coarse_fine):
            self.coarse_fine[i] = cf
        self.num_cf = len(coarse_fine)
        self.tree_boundary = np.ascontiguousarray(tree_boundary, dtype=np.int32).reshape(-1, 2)

    @property
    def num_blocks(self):
        return int(self.level.shape[0])

    @property
    def max_level(self):
        return int(self.level.max()) if self.level.size else 0

    def struct(self) -> Topology:
        t = Topology()
        t.num_blocks = self.num_blocks
        t.n = self.n
        t.halo = self.halo
        t.level = self.level.ctypes.data_as(C.POINTER(C.c_int32))
        t.bounds = self.bounds.ctypes.data_as(C.POINTER(C.c_double))
        t.num_same_level = int(self.same_level.shape[0])
        t.same_level = self.same_level.ctypes.data_as(C.POINTER(C.c_int32))
        t.num_coarse_fine = self.num_cf
        t.coarse_fine = C.cast(self.coarse_fine, C.POINTER(CfFace))
        t.num_tree_boundary = int(self.tree_boundary.shape[0])
        t.tree_boundary = self.tree_boundary.ctypes.data_as(C.POINTER(C.c_int32))
        return t


def make_params(omega, precision=64, handler=HANDLER_CONCRETE, const_rho=0, stats_scope=0,
                bouzidi=0, snapshot=0, specs=(), fast_stats=0):
    arr = (BcSpec * max(1, len(specs)))()
    for i, s in enumerate(specs):
        arr[i] = s
    p = Params()
    p.omega = omega
    p.precision = precision
    p.handler = handler
    p.const_rho = const_rho
    p.stats_scope = stats_scope
    p.bouzidi = bouzidi
    p.snapshot = snapshot
    p.num_specs = len(specs)
    p.specs = C.cast(arr, C.POINTER(BcSpec))
    p.fast_stats = fast_stats
    p._keep = arr  # keep the spec array alive with the struct
    return p


_LIB = None

_I32P = C.POINTER(C.c_int32)
_F64P = C.POINTER(C.c_double)
_U8P = C.POINTER(C.c_uint8)
_VP = C.c_void_p

_SIGNATURES = {
    "octlb_last_error": ([], C.c_char_p),
    "octlb_forest_create": ([_F64P, C.c_int32, C.c_int32, C.c_int32, C.POINTER(_VP)], C.c_int),
    "octlb_forest_destroy": ([_VP], None),
    "octlb_forest_refine_box": ([_VP, _F64P, C.c_int32], C.c_int),
    "octlb_forest_refine_marked": ([_VP, _U8P, C.c_int64], C.c_int),
    "octlb_forest_balance": ([_VP], C.c_int),
    "octlb_forest_num_octants": ([_VP], C.c_int32),
    "octlb_forest_topology": ([_VP, C.c_int32, C.c_int32, C.POINTER(Topology)], C.c_int),
    "octlb_dev_create": ([C.POINTER(Topology), C.POINTER(Params), C.POINTER(_VP)], C.c_int),
    "octlb_dev_destroy": ([_VP], C.c_int),
    "octlb_dev_set_bc_kinds": ([_VP, _U8P], C.c_int),
    "octlb_dev_set_links": ([_VP, C.c_int64, _I32P, _F64P], C.c_int),
    "octlb_dev_init_equilibrium": ([_VP, C.c_double, _F64P, C.c_uint32], C.c_int),
    "octlb_dev_init_field": ([_VP, _F64P, _F64P], C.c_int),
    "octlb_dev_upload_block": ([_VP, C.c_int32, _F64P], C.c_int),
    "octlb_dev_download_block": ([_VP, C.c_int32, _F64P], C.c_int),
    "octlb_dev_download_core": ([_VP, _F64P], C.c_int),
    "octlb_dev_advance": ([_VP, C.c_int32], C.c_int),
    "octlb_dev_advance_graph": ([_VP, C.c_int32], C.c_int),
    "octlb_dev_prepare_graphs": ([_VP], C.c_int),
    "octlb_dev_set_snapshot": ([_VP, C.c_int32], C.c_int),
    "octlb_dev_set_kernel_variant": ([_VP, C.c_int32], C.c_int),
    "octlb_dev_kernel_variant": ([_VP, C.POINTER(C.c_int32), C.POINTER(C.c_int32)], C.c_int),
    "octlb_dev_autotune": ([_VP, C.c_int32, _F64P], C.c_int),
    "octlb_dev_synchronize": ([_VP], C.c_int),
    "octlb_dev_set_step": ([_VP, C.c_int32], C.c_int),
    "octlb_dev_drag": ([_VP, _F64P, _F64P], C.c_int),
    "octlb_dev_mass": ([_VP, _F64P], C.c_int),
    "octlb_dev_average_rho": ([_VP, _F64P], C.c_int),
    "octlb_dev_update_counts": ([_VP, _F64P, _F64P], C.c_int),
    "octlb_dev_kernel_timing": ([_VP, _F64P, C.POINTER(C.c_int64), _F64P], C.c_int),
    "octlb_dev_enable_timing": ([_VP, C.c_int32], C.c_int),
    "octlb_dev_time_kernel": ([_VP, C.c_int32, C.c_int32, _F64P, _F64P], C.c_int),
    "octlb_dev_create_ensemble": ([C.POINTER(Topology), C.c_int32, _I32P, C.POINTER(Params),
                                   C.POINTER(C.c_void_p)], C.c_int),
    "octlb_dev_member_average_rho": ([_VP, C.c_int32, _F64P], C.c_int),
    "octlb_build_links": ([C.c_int32, C.c_int32, _F64P, _U8P, C.c_int64, _F64P, C.c_int64,
                           C.POINTER(C.c_int64), _I32P, _F64P], C.c_int),
    "octlb_build_links_noalloc": ([C.c_int32, C.c_int32, _VP, _VP, C.c_int64, _VP, C.c_int64,
                                   C.POINTER(C.c_int64), _VP, _VP, _VP], C.c_int),
    "octlb_dev_migrate": ([_VP, _VP], C.c_int),
    "octlb_voxelize_part": ([C.c_int32, C.c_int32, _F64P, C.c_int64, _F64P, C.c_int64, _F64P, C.c_int32, _U8P,
                             C.POINTER(C.c_int64)], C.c_int),
    "octlb_dev_macro_fields": ([_VP, _F64P, _F64P], C.c_int),
    "octlb_dev_write_vtk": ([_VP, C.c_int32, C.c_char_p, C.c_char_p], C.c_int),
    "octlb_vtk_write_vts": ([C.c_char_p, C.c_int32, C.c_int32, C.c_int32, _F64P, C.c_int32,
                             C.POINTER(C.c_char_p), _I32P, C.POINTER(_F64P)], C.c_int),
    "octlb_vtk_write_vtm": ([C.c_char_p, C.c_int32, C.POINTER(C.c_char_p)], C.c_int),
    "octlb_vtk_append_pvd": ([C.c_char_p, C.c_int32, C.c_char_p], C.c_int),
    "octlb_adj_create": ([_VP, C.POINTER(C.c_void_p)], C.c_int),
    "octlb_adj_destroy": ([_VP], C.c_int),
    "octlb_adj_run": ([_VP, C.c_int32, C.c_int32, C.c_int32, C.c_int32, _F64P, _F64P, C.c_double, _F64P,
                       C.c_uint32, C.c_int64, _I32P, _F64P, _F64P, _F64P], C.c_int),
    "octlb_adj_final_core": ([_VP, _F64P], C.c_int),
    "octlb_adj_inlet_grad": ([_VP, _F64P], C.c_int),
    "octlb_adj_timing": ([_VP, _F64P, _F64P], C.c_int),
    "octlb_adj_set_tape_budget": ([_VP, C.c_int64, C.POINTER(C.c_int32)], C.c_int),
    "octlb_dev_set_coupling_epilogues": ([_VP, C.c_int32], C.c_int),
    "octlb_dev_coupling_epilogues": ([_VP, C.POINTER(C.c_int32), _F64P], C.c_int),
}

EXPORTED = tuple(_SIGNATURES)


def lib():
    """Load liboctlb_gpu.so; raise if it has not been built."""
    global _LIB
    if _LIB is None:
        if not LIB_PATH.exists():
            raise OctlbError(
                f"{LIB_PATH} is missing: build it with `python -m octlb_amd.build` "
                "(hipcc --offload-arch=gfx950); there is no CPU fallback")
        handle = C.CDLL(str(LIB_PATH), mode=os.RTLD_LOCAL if hasattr(os, "RTLD_LOCAL") else 0)
        for name, (argtypes, restype) in _SIGNATURES.items():
            fn = getattr(handle, name)
            fn.argtypes = argtypes
            fn.restype = restype
        _LIB = handle
    return _LIB


def check(rc: int, what: str = "octlb call"):
    if rc != 0:
        msg = lib().octlb_last_error()
        raise OctlbError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")


def f64p(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_F64P)


def i32p(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_I32P)


def u8p(a: np.ndarray):
    assert a.dtype == np.uint8 and a.flags.c_contiguous
    return a.ctypes.data_as(_U8P)
