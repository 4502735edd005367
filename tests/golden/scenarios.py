"""Single-level scenarios that the reference's own per-block code can run
(oracle/_ref/ref_harness) and that pin the oracle bit for bit.

Inputs come from an integer LCG (not numpy's Generator), so they are
identical on every numpy version; populations are k * 2^-30 with small
integers k, exact in FP64.
"""
from __future__ import annotations

import numpy as np

from octlb_amd import abi, cases
from octlb_amd.forest import single_block_topology

Q = 19


def lcg_values(seed: int, count: int, amplitude: int = 1 << 22) -> np.ndarray:
    """x_{k+1} = a x_k + c mod 2^64 (Knuth MMIX); returns values in
    (-amplitude, amplitude) * 2^-30."""
    a = np.uint64(6364136223846793005)
    c = np.uint64(1442695040888963407)
    x = np.uint64(seed)
    out = np.empty(count, dtype=np.int64)
    with np.errstate(over="ignore"):
        # vectorised jump: compute in chunks with a running state
        state = x
        for i in range(count):
            state = state * a + c
            out[i] = int(state >> np.uint64(33)) % (2 * amplitude) - amplitude
    return out.astype(np.float64) * 2.0 ** -30


def pops_for(case, seed):
    topo = case.topo
    NP = topo.n + 2 * topo.halo
    return lcg_values(seed, topo.num_blocks * NP ** 3 * Q).reshape(topo.num_blocks, NP, NP, NP, Q)


def _block_case(handler=abi.HANDLER_CONCRETE, bouzidi=1, const_rho=0):
    n, h = 8, 1
    topo = single_block_topology(n, h)
    kinds = np.zeros((1, n, n, n), np.uint8)
    specs = [abi.make_spec(abi.XMIN, abi.BC_INTERP_VELOCITY, inlet=abi.INLET_UNIFORM,
                           inlet_u=(0.03, 0.0, 0.0), ramp_end_t=4.0),
             abi.make_spec(abi.XMAX, abi.BC_INTERP_PRESSURE, rho_target=1.0)]
    cases.stamp_tree_boundary(kinds, topo.tree_boundary, specs)
    kinds[0, 3:5, 3:5, 3:5] = abi.KIND_SOLID
    kinds[0, 3:5, 3:5, 2] = abi.KIND_BOUZIDI
    kinds[0, 6, 1, 1] = abi.KIND_BOUNCE_BACK
    kinds[0, 5, 6, 6] = abi.KIND_MOVING_BOUNCE_BACK
    kinds[0, 2, 6, 6] = abi.KIND_OUTFLOW
    # pull sources x - c_i that are solid-like for cells around the obstacle
    link, q = [], []
    qs = [0.2, 0.35, 0.5, 0.7, 0.95, 1e-12, 0.4999999999999999]
    k = 0
    for x in range(2, 6):
        for y in range(2, 6):
            for z in range(0, 7):
                if kinds[0, x, y, z] not in (abi.KIND_BULK,):
                    continue
                for i in range(1, Q):
                    sx, sy, sz = x - cases.C[i, 0], y - cases.C[i, 1], z - cases.C[i, 2]
                    if not (0 <= sx < n and 0 <= sy < n and 0 <= sz < n):
                        continue
                    if kinds[0, sx, sy, sz] in (abi.KIND_SOLID, abi.KIND_BOUZIDI):
                        link.append([0, (x * n + y) * n + z, i])
                        q.append(qs[k % len(qs)])
                        k += 1
    case = cases.Case(name="block", topo=topo, omega=1 / 0.6, handler=handler,
                      const_rho=const_rho, bouzidi=bouzidi, snapshot=1, specs=specs, kinds=kinds,
                      link=np.array(link, np.int32).reshape(-1, 3), q=np.array(q),
                      drag_bbox=np.array([0.2, 0.2, 0.0, 0.8, 0.8, 1.0]))
    if not bouzidi:
        case.link = np.zeros((0, 3), np.int32)
        case.q = np.zeros(0)
    return case


def _inlet_profiles_case():
    n, h = 8, 1
    topo = single_block_topology(n, h, edge=0.5)
    specs = [abi.make_spec(abi.XMIN, abi.BC_INTERP_VELOCITY, inlet=abi.INLET_POISEUILLE,
                           inlet_u=(1.0, 0.0, 0.0), u_peak=0.04, n_inplane=(8, 8), ramp_end_t=6.0),
             abi.make_spec(abi.YMIN, abi.BC_INTERP_VELOCITY, inlet=abi.INLET_CHANNEL,
                           inlet_u=(0.0, 1.0, 0.0), u_peak=0.02, lo0=0.05, hi0=0.45, lo1=0.1,
                           hi1=0.4, ramp_end_t=0.0),
             abi.make_spec(abi.XMAX, abi.BC_INTERP_PRESSURE, rho_target=1.001),
             abi.make_spec(abi.ZMAX, abi.BC_INTERP_VELOCITY, u_wall=(0.01, 0.0, 0.0))]
    kinds = np.zeros((1, n, n, n), np.uint8)
    cases.stamp_tree_boundary(kinds, topo.tree_boundary, specs)
    return cases.Case(name="inlets", topo=topo, omega=1 / 0.7, handler=abi.HANDLER_CONCRETE,
                      const_rho=0, bouzidi=1, snapshot=0, specs=specs, kinds=kinds)


def _multiblock_case():
    return cases.cylinder_channel(bricks_x=4, bricks_yz=2, base_level=0, fine=False, xc=0.5,
                                  radius=0.15)


def scenarios():
    """name -> (case factory, seed, steps, t0)."""
    return {
        "block_concrete_bouzidi": (lambda: _block_case(), 101, 6, 0),
        "block_noop_rotate": (lambda: _block_case(abi.HANDLER_NOOP, bouzidi=0), 102, 4, 0),
        "block_noop_const_rho": (lambda: _block_case(abi.HANDLER_NOOP, bouzidi=0, const_rho=1), 103, 4, 0),
        "block_concrete_const_rho_rotate": (lambda: _block_case(abi.HANDLER_CONCRETE, bouzidi=0, const_rho=1), 104, 5, 0),
        "cavity10_h3": (lambda: cases.cavity(10, tau=0.6, u_lid=0.05), 105, 6, 0),
        "inlet_profiles": (lambda: _inlet_profiles_case(), 106, 5, 3),
        "multiblock_cylinder": (lambda: _multiblock_case(), 107, 5, 0),
    }
