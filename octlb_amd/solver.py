"""Device solver handle over liboctlb_gpu.so.

``GpuSolver`` stands where the reference constructs BlockCollection +
GhostSchedule + LevelCoupler + DomainBoundaryHandler + TimeLoop
(examples/*/..._case.h) and calls ``TimeLoop::advance_one`` (time_loop.h:56).
Every call goes to the HIP library; there is no CPU path.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi


class GpuSolver:
    def __init__(self, topo: abi.TopologyArrays, params: abi.Params):
        self._lib = abi.lib()
        self.topo = topo
        self.params = params
        self.n = topo.n
        self.halo = topo.halo
        self.nb = topo.num_blocks
        self.ncore = self.n ** 3
        self.np_ = self.n + 2 * self.halo
        h = C.c_void_p()
        t = topo.struct()
        abi.check(self._lib.octlb_dev_create(C.byref(t), C.byref(params), C.byref(h)), "dev create")
        self._h = h

    @classmethod
    def ensemble(cls, ens, precision: int = 64):
        """Many independent cases in one device state (``cases.Ensemble``),
        stepped by the same launches: octlb_dev_create_ensemble."""
        self = cls.__new__(cls)
        self._lib = abi.lib()
        self.topo = ens.topo
        params_list = ens.params_for(precision)
        self.params = params_list[0]
        self.n = ens.topo.n
        self.halo = ens.topo.halo
        self.nb = ens.topo.num_blocks
        self.ncore = self.n ** 3
        self.np_ = self.n + 2 * self.halo
        self.members = len(params_list)
        begin = np.ascontiguousarray(ens.member_begin, dtype=np.int32)
        plist = (abi.Params * self.members)(*params_list)
        h = C.c_void_p()
        t = ens.topo.struct()
        abi.check(self._lib.octlb_dev_create_ensemble(C.byref(t), self.members, abi.i32p(begin), plist, C.byref(h)),
                  "ensemble create")
        self._h = h
        return self

    def member_average_rho(self, m):
        out = np.zeros(1)
        abi.check(self._lib.octlb_dev_member_average_rho(self._h, int(m), abi.f64p(out)), "member_average_rho")
        return float(out[0])

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.octlb_dev_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- setup
    def set_bc_kinds(self, kinds):
        k = np.ascontiguousarray(kinds, dtype=np.uint8).reshape(-1)
        assert k.size == self.nb * self.ncore
        abi.check(self._lib.octlb_dev_set_bc_kinds(self._h, abi.u8p(k)), "set_bc_kinds")

    def set_links(self, link, q):
        link = np.ascontiguousarray(link, dtype=np.int32).reshape(-1, 3)
        q = np.ascontiguousarray(q, dtype=np.float64).reshape(-1)
        assert link.shape[0] == q.size
        abi.check(self._lib.octlb_dev_set_links(self._h, int(q.size), abi.i32p(link), abi.f64p(q)),
                  "set_links")

    def init_equilibrium(self, rho0=1.0, u0=(0.0, 0.0, 0.0), kind_mask=0xFF):
        u = np.ascontiguousarray(u0, dtype=np.float64)
        abi.check(self._lib.octlb_dev_init_equilibrium(self._h, float(rho0), abi.f64p(u),
                                                       int(kind_mask)), "init")

    def init_field(self, rho, u):
        r = np.ascontiguousarray(rho, dtype=np.float64).reshape(-1)
        uu = np.ascontiguousarray(u, dtype=np.float64).reshape(-1)
        abi.check(self._lib.octlb_dev_init_field(self._h, abi.f64p(r), abi.f64p(uu)), "init_field")

    def upload_block(self, b, aos):
        a = np.ascontiguousarray(aos, dtype=np.float64).reshape(-1)
        assert a.size == self.np_ ** 3 * abi.Q
        abi.check(self._lib.octlb_dev_upload_block(self._h, int(b), abi.f64p(a)), "upload")

    def download_block(self, b):
        a = np.empty(self.np_ ** 3 * abi.Q)
        abi.check(self._lib.octlb_dev_download_block(self._h, int(b), abi.f64p(a)), "download")
        return a.reshape(self.np_, self.np_, self.np_, abi.Q)

    def download_core(self):
        a = np.empty(self.nb * self.ncore * abi.Q)
        abi.check(self._lib.octlb_dev_download_core(self._h, abi.f64p(a)), "download_core")
        return a.reshape(self.nb, self.n, self.n, self.n, abi.Q)

    def migrate_from(self, other):
        """Take over another handle's state after a forest change
        (octlb_dev_migrate): unchanged blocks copied, refined blocks
        prolonged, coarsened blocks restricted."""
        abi.check(self._lib.octlb_dev_migrate(self._h, other._h), "migrate")

    # ---- output (the reference's VTK samplers and writers)
    def macro_fields(self):
        """(rho [nb, n, n, n], u [nb, n, n, n, 3]) per core cell in VTK cell
        order [block][k][j][i] (x fastest)."""
        rho = np.empty(self.nb * self.ncore)
        u = np.empty(self.nb * self.ncore * 3)
        abi.check(self._lib.octlb_dev_macro_fields(self._h, abi.f64p(rho), abi.f64p(u)), "macro_fields")
        return rho.reshape(self.nb, self.n, self.n, self.n), u.reshape(self.nb, self.n, self.n, self.n, 3)

    def write_vtk(self, it, output_dir, base_name="cylinder3d"):
        """Cylinder3dCase::write_vtk_timestep: one .vts per block (velocity,
        pressure), the .vtm index and an entry in base_name.pvd."""
        abi.check(self._lib.octlb_dev_write_vtk(self._h, int(it), str(output_dir).encode(), base_name.encode()),
                  "write_vtk")

    # ---- stepping
    def advance(self, steps=1):
        abi.check(self._lib.octlb_dev_advance(self._h, int(steps)), "advance")

    def advance_graph(self, steps=1):
        abi.check(self._lib.octlb_dev_advance_graph(self._h, int(steps)), "advance_graph")

    def set_snapshot(self, on=True):
        """TimeLoop::set_snapshot_post_collide: snapshot the drag links'
        post-collide populations in every following step."""
        abi.check(self._lib.octlb_dev_set_snapshot(self._h, 1 if on else 0), "set_snapshot")

    def set_kernel_variant(self, variant):
        """0: every block pulls through its LDS box; 1: plain interior blocks
        push, the rest pull (default)."""
        abi.check(self._lib.octlb_dev_set_kernel_variant(self._h, int(variant)), "set_kernel_variant")

    def set_coupling_epilogues(self, mode):
        """Bit 0: the half-time prolongation runs in the preceding update's
        launch; bit 1: the restriction runs in the finest level's second
        update launch (default 3; the autotuner decides per workload)."""
        abi.check(self._lib.octlb_dev_set_coupling_epilogues(self._h, int(mode)), "set_coupling_epilogues")

    def coupling_epilogues(self):
        """(mode in effect, autotuner's [ms without, ms with] per coarse step)."""
        m = C.c_int32()
        ms = np.zeros(2)
        abi.check(self._lib.octlb_dev_coupling_epilogues(self._h, C.byref(m), abi.f64p(ms)), "coupling_epilogues")
        return int(m.value), ms

    def kernel_variant(self):
        """(variant in use, blocks updated by push)."""
        v, n = C.c_int32(), C.c_int32()
        abi.check(self._lib.octlb_dev_kernel_variant(self._h, C.byref(v), C.byref(n)), "kernel_variant")
        return int(v.value), int(n.value)

    def autotune(self, reps=20):
        """Pick the faster kernel variant on the current state; returns
        (chosen variant, per-variant fused-kernel ms per coarse step)."""
        ms = np.zeros(2)
        abi.check(self._lib.octlb_dev_autotune(self._h, int(reps), abi.f64p(ms)), "autotune")
        return self.kernel_variant()[0], ms

    def prepare_graphs(self):
        """Capture every graph advance_graph replays, without running it."""
        abi.check(self._lib.octlb_dev_prepare_graphs(self._h), "prepare_graphs")

    def synchronize(self):
        abi.check(self._lib.octlb_dev_synchronize(self._h), "synchronize")

    def set_step(self, step):
        abi.check(self._lib.octlb_dev_set_step(self._h, int(step)), "set_step")

    # ---- observables
    def drag(self, bbox):
        b = np.ascontiguousarray(bbox, dtype=np.float64)
        out = np.zeros(3)
        abi.check(self._lib.octlb_dev_drag(self._h, abi.f64p(b), abi.f64p(out)), "drag")
        return out

    def mass(self):
        out = np.zeros(1)
        abi.check(self._lib.octlb_dev_mass(self._h, abi.f64p(out)), "mass")
        return float(out[0])

    def average_rho(self):
        out = np.zeros(1)
        abi.check(self._lib.octlb_dev_average_rho(self._h, abi.f64p(out)), "average_rho")
        return float(out[0])

    def update_counts(self):
        a = np.zeros(1)
        f = np.zeros(1)
        abi.check(self._lib.octlb_dev_update_counts(self._h, abi.f64p(a), abi.f64p(f)), "counts")
        return float(a[0]), float(f[0])

    def time_kernel(self, level, reps=50):
        """Mean fused-kernel launch duration (ms) of one level and the cells
        one launch updates; leaves the simulation state unchanged."""
        ms, cells = C.c_double(), C.c_double()
        abi.check(self._lib.octlb_dev_time_kernel(self._h, int(level), int(reps), C.byref(ms),
                                                  C.byref(cells)), "time_kernel")
        return ms.value, cells.value

    def enable_timing(self, on=True):
        abi.check(self._lib.octlb_dev_enable_timing(self._h, 1 if on else 0), "timing")

    def kernel_timing(self):
        ms = C.c_double()
        launches = C.c_int64()
        upd = C.c_double()
        abi.check(self._lib.octlb_dev_kernel_timing(self._h, C.byref(ms), C.byref(launches),
                                                    C.byref(upd)), "timing")
        return ms.value, launches.value, upd.value
