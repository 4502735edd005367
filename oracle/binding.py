"""ctypes binding of the CPU restatement (liboctlb_oracle.so) and of the
reference harness (oracle/_ref/ref_harness).  TEST INFRASTRUCTURE: used by
tests/, the smoke check and bench.py's cpu_baseline leg only.

``OracleSim`` exposes the same calls as ``octlb_amd.solver.GpuSolver`` so a
test can drive both with one case description.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import tempfile
from pathlib import Path

import numpy as np

from octlb_amd import abi

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboctlb_oracle.so"
REF_HARNESS = HERE / "_ref" / "ref_harness"
Q = 19

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            raise RuntimeError(f"{LIB} missing: run `make -C oracle`")
        h = C.CDLL(str(LIB))
        vp, i32p, f64p, u8p = C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_double), C.POINTER(C.c_uint8)
        sig = {
            "ol_create": ([C.POINTER(abi.Topology), C.POINTER(abi.Params)], vp),
            "ol_destroy": ([vp], None),
            "ol_last_error": ([], C.c_char_p),
            "ol_set_bc_kinds": ([vp, u8p], C.c_int),
            "ol_set_links": ([vp, C.c_int64, i32p, f64p], C.c_int),
            "ol_init_equilibrium": ([vp, C.c_double, f64p, C.c_uint32], C.c_int),
            "ol_init_field": ([vp, f64p, f64p], C.c_int),
            "ol_set_block": ([vp, C.c_int32, f64p], C.c_int),
            "ol_get_block": ([vp, C.c_int32, f64p], C.c_int),
            "ol_get_core": ([vp, f64p], C.c_int),
            "ol_get_snapshot_block": ([vp, C.c_int32, f64p], C.c_int),
            "ol_advance": ([vp, C.c_int32], C.c_int),
            "ol_set_threads": ([vp, C.c_int32], C.c_int),
            "ol_set_step": ([vp, C.c_int32], C.c_int),
            "ol_drag": ([vp, f64p, f64p], C.c_int),
            "ol_mass": ([vp], C.c_double),
            "ol_average_rho": ([vp], C.c_double),
            "ol_counters": ([vp, i32p, i32p, C.c_int32, i32p, C.c_int32, i32p], C.c_int),
            "ol_plan_size": ([vp], C.c_int32),
            "ol_plan": ([vp, i32p], C.c_int),
            "ol_num_edges": ([vp], C.c_int32),
            "ol_edges": ([vp, i32p], C.c_int),
            "ol_block_collide": ([vp, C.c_int32, C.c_double], C.c_int),
            "ol_block_stream": ([vp, C.c_int32], C.c_int),
            "ol_block_post_stream": ([vp, C.c_int32, C.c_double], C.c_int),
            "ol_coupler_init": ([vp], C.c_int),
            "ol_apply_prolongation": ([vp, C.c_int32, C.c_int32], C.c_int),
            "ol_apply_restriction": ([vp, C.c_int32], C.c_int),
            "ol_build_links": ([C.c_int32, C.c_int32, f64p, u8p, C.c_int64, f64p, C.c_int64, i32p, f64p],
                               C.c_int64),
            "ol_migrate_block": ([C.c_int32, C.c_int32, C.c_int32, i32p, C.POINTER(f64p), C.c_double, f64p],
                                 C.c_int),
            "ol_voxelize_part": ([C.c_int32, C.c_int32, f64p, C.c_int64, f64p, C.c_int64, f64p, C.c_int32, u8p],
                                 C.c_int),
            "ol_adjoint": ([vp, C.c_int32, C.c_int32, C.c_int32, f64p, f64p, C.c_double, f64p, C.c_uint32,
                            C.c_int64, i32p, f64p, f64p, f64p], C.c_int),
        }
        for name, (a, r) in sig.items():
            fn = getattr(h, name)
            fn.argtypes = a
            fn.restype = r
        _lib = h
    return _lib


def _ok(rc, what):
    if rc != 0:
        raise RuntimeError(f"oracle {what}: {lib().ol_last_error().decode()}")


class OracleSim:
    def __init__(self, topo: abi.TopologyArrays, params: abi.Params):
        self.topo = topo
        self.params = params
        self.n, self.halo, self.nb = topo.n, topo.halo, topo.num_blocks
        self.np_ = self.n + 2 * self.halo
        self.ncore = self.n ** 3
        t = topo.struct()
        self._h = lib().ol_create(C.byref(t), C.byref(params))
        if not self._h:
            raise RuntimeError(f"oracle create: {lib().ol_last_error().decode()}")

    def close(self):
        if getattr(self, "_h", None):
            lib().ol_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_threads(self, n):
        _ok(lib().ol_set_threads(self._h, int(n)), "threads")

    def set_bc_kinds(self, kinds):
        k = np.ascontiguousarray(kinds, dtype=np.uint8).reshape(-1)
        _ok(lib().ol_set_bc_kinds(self._h, abi.u8p(k)), "kinds")

    def set_links(self, link, q):
        link = np.ascontiguousarray(link, dtype=np.int32).reshape(-1, 3)
        q = np.ascontiguousarray(q, dtype=np.float64).reshape(-1)
        _ok(lib().ol_set_links(self._h, int(q.size), abi.i32p(link), abi.f64p(q)), "links")

    def init_equilibrium(self, rho0=1.0, u0=(0.0, 0.0, 0.0), kind_mask=0xFF):
        u = np.ascontiguousarray(u0, dtype=np.float64)
        _ok(lib().ol_init_equilibrium(self._h, float(rho0), abi.f64p(u), int(kind_mask)), "init")

    def init_field(self, rho, u):
        r = np.ascontiguousarray(rho, dtype=np.float64).reshape(-1)
        uu = np.ascontiguousarray(u, dtype=np.float64).reshape(-1)
        _ok(lib().ol_init_field(self._h, abi.f64p(r), abi.f64p(uu)), "init_field")

    def upload_block(self, b, aos):
        a = np.ascontiguousarray(aos, dtype=np.float64).reshape(-1)
        _ok(lib().ol_set_block(self._h, int(b), abi.f64p(a)), "set_block")

    def download_block(self, b):
        a = np.empty(self.np_ ** 3 * Q)
        _ok(lib().ol_get_block(self._h, int(b), abi.f64p(a)), "get_block")
        return a.reshape(self.np_, self.np_, self.np_, Q)

    def download_core(self):
        a = np.empty(self.nb * self.ncore * Q)
        _ok(lib().ol_get_core(self._h, abi.f64p(a)), "get_core")
        return a.reshape(self.nb, self.n, self.n, self.n, Q)

    def advance(self, steps=1):
        _ok(lib().ol_advance(self._h, int(steps)), "advance")

    def set_step(self, step):
        _ok(lib().ol_set_step(self._h, int(step)), "set_step")

    def drag(self, bbox):
        b = np.ascontiguousarray(bbox, dtype=np.float64)
        out = np.zeros(3)
        _ok(lib().ol_drag(self._h, abi.f64p(b), abi.f64p(out)), "drag")
        return out

    def adjoint(self, steps, window, direction, bbox, u0, rho0=1.0, kind_mask=1, link=None):
        """J = mean over steps window[0]+1 .. window[1] of direction . drag(bbox)
        from feq(rho0, u0); returns (J, dJ/du0, dJ/dq per link row)."""
        d = np.ascontiguousarray(direction, dtype=np.float64)
        bb = np.ascontiguousarray(bbox, dtype=np.float64)
        u = np.ascontiguousarray(u0, dtype=np.float64)
        link = np.zeros((0, 3), np.int32) if link is None else np.ascontiguousarray(link, dtype=np.int32).reshape(-1, 3)
        J = np.zeros(1)
        du = np.zeros(3)
        dq = np.zeros(max(len(link), 1))
        _ok(lib().ol_adjoint(self._h, int(steps), int(window[0]), int(window[1]), abi.f64p(d), abi.f64p(bb),
                             float(rho0), abi.f64p(u), int(kind_mask), int(len(link)), abi.i32p(link),
                             abi.f64p(J), abi.f64p(du), abi.f64p(dq)), "adjoint")
        return float(J[0]), du, dq[: len(link)]

    def mass(self):
        return float(lib().ol_mass(self._h))

    def average_rho(self):
        return float(lib().ol_average_rho(self._h))

    def counters(self, max_levels=8, max_calls=4096):
        col = np.zeros(max_levels, np.int32)
        st = np.zeros(max_levels, np.int32)
        calls = np.zeros(2 * max_calls, np.int32)
        nc = np.zeros(1, np.int32)
        _ok(lib().ol_counters(self._h, abi.i32p(col), abi.i32p(st), max_levels, abi.i32p(calls),
                              max_calls, abi.i32p(nc)), "counters")
        return col, st, calls[: 2 * int(nc[0])].reshape(-1, 2)

    def plan(self):
        m = int(lib().ol_plan_size(self._h))
        rows = np.zeros(max(m, 1) * 12, np.int32)
        _ok(lib().ol_plan(self._h, abi.i32p(rows)), "plan")
        return rows[: m * 12].reshape(m, 12)

    def edges(self):
        m = int(lib().ol_num_edges(self._h))
        rows = np.zeros(max(m, 1) * 4, np.int32)
        _ok(lib().ol_edges(self._h, abi.i32p(rows)), "edges")
        return rows[: m * 4].reshape(m, 4)

    def block_collide(self, b, t=0.0):
        _ok(lib().ol_block_collide(self._h, int(b), float(t)), "collide")

    def block_stream(self, b):
        _ok(lib().ol_block_stream(self._h, int(b)), "stream")

    def block_post_stream(self, b, t=0.0):
        _ok(lib().ol_block_post_stream(self._h, int(b), float(t)), "post_stream")

    def coupler_init(self):
        _ok(lib().ol_coupler_init(self._h), "coupler_init")

    def apply_prolongation(self, level, half_time):
        _ok(lib().ol_apply_prolongation(self._h, int(level), 1 if half_time else 0), "prolongation")

    def apply_restriction(self, fine_level):
        _ok(lib().ol_apply_restriction(self._h, int(fine_level)), "restriction")


def build_links(bounds, n, material, triangles):
    """BouzidiLinkData::Build restated (links (block, cell, pop) int32 [L,3],
    q float64 [L]); material in MaterialField order [block][k][j][i]."""
    h = lib()
    bounds = np.ascontiguousarray(bounds, np.float64).reshape(-1, 6)
    material = np.ascontiguousarray(material, np.uint8)
    tri = np.ascontiguousarray(triangles, np.float64).reshape(-1, 9)
    nb = bounds.shape[0]
    f64p, u8p, i32p = C.POINTER(C.c_double), C.POINTER(C.c_uint8), C.POINTER(C.c_int32)
    args = (nb, n, bounds.ctypes.data_as(f64p), material.ctypes.data_as(u8p), tri.shape[0],
            tri.ctypes.data_as(f64p) if tri.size else None)
    cnt = h.ol_build_links(*args, 0, None, None)
    if cnt < 0:
        raise RuntimeError(h.ol_last_error().decode())
    link = np.zeros((max(cnt, 1), 3), np.int32)
    q = np.zeros(max(cnt, 1))
    h.ol_build_links(*args, cnt, link.ctypes.data_as(i32p), q.ctypes.data_as(f64p))
    return link[:cnt], q[:cnt]


def migrate_block(mode, n, h, oct, views, tau):
    """ol_migrate_block: the halo-inclusive view [(n+2h)^3, 19] of one block
    of an adapted forest.  mode 0 copy / 1 refine (views = [parent], oct =
    octant) / 2 coarsen (views = the eight children in octant order)."""
    E = n + 2 * h
    vs = [np.ascontiguousarray(v, np.float64).reshape(E ** 3, Q) for v in views]
    f64p = C.POINTER(C.c_double)
    arr = (f64p * 8)(*([v.ctypes.data_as(f64p) for v in vs] + [None] * (8 - len(vs))))
    oc = np.ascontiguousarray(oct, np.int32)
    out = np.zeros((E ** 3, Q))
    _ok(lib().ol_migrate_block(int(mode), int(n), int(h), oc.ctypes.data_as(C.POINTER(C.c_int32)), arr,
                               float(tau), out.ctypes.data_as(f64p)), "migrate_block")
    return out


def voxelize_part(bounds, n, soup, role, inner=None):
    """ol_voxelize_part: MaterialKind bytes [block][k][j][i] of one part
    (role 0 internal channel, 1 external obstacle)."""
    bounds = np.ascontiguousarray(bounds, np.float64).reshape(-1, 6)
    tri = np.ascontiguousarray(soup, np.float64).reshape(-1, 9)
    inn = np.ascontiguousarray(inner if inner is not None else np.zeros((0, 9)), np.float64).reshape(-1, 9)
    nb = bounds.shape[0]
    out = np.zeros((nb, n, n, n), np.uint8)
    f64p, u8p = C.POINTER(C.c_double), C.POINTER(C.c_uint8)
    _ok(lib().ol_voxelize_part(nb, n, bounds.ctypes.data_as(f64p), tri.shape[0],
                               tri.ctypes.data_as(f64p) if tri.size else None, inn.shape[0],
                               inn.ctypes.data_as(f64p) if inn.size else None, int(role),
                               out.ctypes.data_as(u8p)), "voxelize_part")
    return out


def setup(sim, case, precision_note=None):
    """Apply a ``octlb_amd.cases.Case`` to a GpuSolver or OracleSim."""
    return case.apply(sim)


# ------------------------------------------------------------ reference harness
def ref_available() -> bool:
    return REF_HARNESS.exists() and os.access(REF_HARNESS, os.X_OK)


def run_reference(case, pops, steps, t0=0):
    """Run the reference's own per-block solver code on a single-level case.

    pops: [nb, NP, NP, NP, 19] initial populations.  Returns
    (pops_out, average_rho, mass, drag[3])."""
    topo = case.topo
    if topo.max_level != 0:
        raise ValueError("the reference harness drives single-level forests only")
    nb, n, h = topo.num_blocks, topo.n, topo.halo
    specs = (abi.BcSpec * max(1, len(case.specs)))(*case.specs)
    with tempfile.TemporaryDirectory() as td:
        scen = Path(td) / "scen.bin"
        outp = Path(td) / "out.bin"
        with open(scen, "wb") as fh:
            hdr = np.array([0x4f43544c, n, h, nb, case.handler, case.const_rho, case.stats_scope,
                            case.bouzidi, case.snapshot, steps, len(case.specs),
                            topo.same_level.shape[0], case.link.shape[0],
                            topo.tree_boundary.shape[0], t0, 0], dtype=np.int32)
            fh.write(hdr.tobytes())
            fh.write(np.array([case.omega], dtype=np.float64).tobytes())
            fh.write(bytes(specs)[: C.sizeof(abi.BcSpec) * len(case.specs)])
            fh.write(np.ascontiguousarray(topo.bounds, dtype=np.float64).tobytes())
            fh.write(np.ascontiguousarray(topo.same_level, dtype=np.int32).tobytes())
            fh.write(np.ascontiguousarray(topo.tree_boundary, dtype=np.int32).tobytes())
            fh.write(np.ascontiguousarray(case.kinds, dtype=np.uint8).tobytes())
            fh.write(np.ascontiguousarray(case.link, dtype=np.int32).tobytes())
            fh.write(np.ascontiguousarray(case.q, dtype=np.float64).tobytes())
            fh.write(np.ascontiguousarray(pops, dtype=np.float64).tobytes())
            bbox = case.drag_bbox if case.drag_bbox is not None else np.array([-1e30] * 3 + [1e30] * 3)
            fh.write(np.ascontiguousarray(bbox, dtype=np.float64).tobytes())
        subprocess.run([str(REF_HARNESS), str(scen), str(outp)], check=True)
        raw = np.fromfile(outp, dtype=np.float64)
    NP = n + 2 * h
    m = nb * NP ** 3 * Q
    return raw[:m].reshape(nb, NP, NP, NP, Q), raw[m], raw[m + 1], raw[m + 2:m + 5]
