"""Gradients of a time-averaged drag objective through the device step.

``DragAdjoint`` drives octlb_adj_create / octlb_adj_run: J = mean over the
steps window[0]+1 .. window[1] of direction . drag(bbox), the momentum-
exchange force of MomentumExchangeDrag::force_on_fluid_if
(drag_postprocessor.h:66-115) summed like octlb_dev_drag, from feq(rho0, u0)
on the bulk cells; gradients with respect to u0 and every Bouzidi wall
fraction q (bouzidi_pull.h:8-26).  ``DragObjective`` wraps it as a
torch.autograd.Function so that u0 and q can sit in a torch graph.

The reference has no adjoint; the gradients are defined by the forward this
library runs (concrete handler, BGK, bounce-back and Bouzidi cells; with
several levels the Lagrava coupler, with velocity / pressure boundaries the
FD pass and the velocity-Dirichlet collide, and then also the gradient with
respect to the inlet velocity).  Everything runs in the HIP library; there is
no CPU path.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import abi
from .solver import GpuSolver


class DragAdjoint:
    def __init__(self, case, steps: int, window: tuple, checkpoint: int = 100, direction=(1.0, 0.0, 0.0),
                 rho0: float = 1.0, precision: int = 64, tape_budget: int = -1):
        self.case = case
        self.steps = int(steps)
        self.window = (int(window[0]), int(window[1]))
        self.checkpoint = int(checkpoint)
        self.direction = np.ascontiguousarray(direction, dtype=np.float64)
        self.rho0 = float(rho0)
        self.precision = int(precision)  # forward states: 64, or 32 with the adjoint still in FP64
        self.timing = (0.0, 0.0)
        self.inlet_grad = np.zeros(3)
        # multi-level engine: bytes for one segment's tapes (-1: automatic,
        # 0: always recompute); ``packed`` tells which mode the last run used
        self.tape_budget = int(tape_budget)
        self.packed = False

    def _params(self, u_in):
        p = self.case.params(self.precision)
        if u_in is not None:
            for k in range(p.num_specs):
                if p.specs[k].inlet != abi.INLET_NONE:
                    for d in range(3):
                        p.specs[k].inlet_u[d] = float(u_in[d])
        return p

    def _handles(self, q, u_in=None):
        sol = GpuSolver(self.case.topo, self._params(u_in))
        sol.set_bc_kinds(self.case.kinds)
        if len(self.case.link):
            sol.set_links(self.case.link, q)
        h = C.c_void_p()
        abi.check(abi.lib().octlb_adj_create(sol._h, C.byref(h)), "adjoint create")
        return sol, h

    def evaluate(self, u0=None, q=None, final_core: bool = False, u_in=None):
        """(J, dJ/du0 [3], dJ/dq [links]) and, if asked, the core populations
        after the forward sweep.  u_in replaces the inlet velocity vector of
        the case's inlet field; dJ/du_in of the run is in ``inlet_grad``."""
        lib = abi.lib()
        u0 = np.ascontiguousarray(self.case.init_u0 if u0 is None else u0, dtype=np.float64)
        q = np.ascontiguousarray(self.case.q if q is None else q, dtype=np.float64)
        link = np.ascontiguousarray(self.case.link, dtype=np.int32).reshape(-1, 3)
        sol, h = self._handles(q, u_in)
        try:
            abi.check(lib.octlb_adj_set_tape_budget(h, self.tape_budget, None), "adjoint tape budget")
            J = np.zeros(1)
            du = np.zeros(3)
            dq = np.zeros(max(len(link), 1))
            bbox = np.ascontiguousarray(self.case.drag_bbox, dtype=np.float64)
            abi.check(lib.octlb_adj_run(h, self.steps, self.window[0], self.window[1], self.checkpoint,
                                        abi.f64p(self.direction), abi.f64p(bbox), self.rho0, abi.f64p(u0),
                                        int(self.case.init_mask), len(link), abi.i32p(link), abi.f64p(J),
                                        abi.f64p(du), abi.f64p(dq)), "adjoint run")
            fw, bw = np.zeros(1), np.zeros(1)
            abi.check(lib.octlb_adj_timing(h, abi.f64p(fw), abi.f64p(bw)), "adjoint timing")
            gi = np.zeros(3)
            abi.check(lib.octlb_adj_inlet_grad(h, abi.f64p(gi)), "adjoint inlet gradient")
            self.inlet_grad = gi
            self.timing = (float(fw[0]), float(bw[0]))
            packed = C.c_int32(0)
            abi.check(lib.octlb_adj_set_tape_budget(h, self.tape_budget, C.byref(packed)), "adjoint tape budget")
            self.packed = bool(packed.value)
            core = None
            if final_core:
                core = np.empty(sol.nb * sol.ncore * 19)
                abi.check(lib.octlb_adj_final_core(h, abi.f64p(core)), "adjoint state")
                core = core.reshape(sol.nb, sol.n, sol.n, sol.n, 19)
        finally:
            lib.octlb_adj_destroy(h)
            sol.close()
        out = (float(J[0]), du, dq[: len(link)])
        return out + (core,) if final_core else out


class DragObjective(torch.autograd.Function):
    """J(u0, q) with its gradient from the device adjoint:
    ``J = DragObjective.apply(u0, q, adjoint)`` for float64 tensors u0 [3]
    and q [links]."""

    @staticmethod
    def forward(ctx, u0, q, adjoint: DragAdjoint):
        J, du, dq = adjoint.evaluate(u0.detach().cpu().double().numpy(), q.detach().cpu().double().numpy())
        ctx.save_for_backward(torch.from_numpy(du).to(u0), torch.from_numpy(np.ascontiguousarray(dq)).to(q))
        return torch.tensor(J, dtype=torch.float64, device=u0.device)

    @staticmethod
    def backward(ctx, grad):
        du, dq = ctx.saved_tensors
        return grad * du, grad * dq, None
