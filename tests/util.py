"""Shared helpers for the parity tests: run one case on the GPU library and
on the CPU oracle with identical inputs."""
from __future__ import annotations

import numpy as np

from octlb_amd import abi
from oracle.binding import OracleSim, setup

Q = 19


def random_pops(case, seed, scale=0.01):
    topo = case.topo
    NP = topo.n + 2 * topo.halo
    rng = np.random.default_rng(seed)
    return rng.normal(0.0, scale, size=(topo.num_blocks, NP, NP, NP, Q))


def make_oracle(case, pops=None):
    sim = OracleSim(case.topo, case.params(64))
    setup(sim, case)
    if pops is not None:
        for b in range(case.topo.num_blocks):
            sim.upload_block(b, pops[b])
    return sim


def make_gpu(case, precision=64, pops=None):
    from octlb_amd.solver import GpuSolver
    sim = GpuSolver(case.topo, case.params(precision))
    setup(sim, case)
    if pops is not None:
        for b in range(case.topo.num_blocks):
            sim.upload_block(b, pops[b])
    return sim


def rel_err(a, b):
    den = max(np.abs(b).max(), 1e-300)
    return float(np.abs(a - b).max() / den)
