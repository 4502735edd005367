"""Brick forest (host) -- the p4est-free OctreeForest/FacePairList provider.

Mirrors the calls the reference cases make (octree_forest.cpp:204-282,
face_pair_list.cpp:480-570): construct over a domain with a brick of trees,
refine by bounds predicate / by marks, 2:1 face balance, then flatten into a
topology (levels, bounds, same-level / coarse-fine / tree-boundary faces).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi


class Forest:
    def __init__(self, domain, bricks=(1, 1, 1)):
        self._lib = abi.lib()
        dom = np.ascontiguousarray(domain, dtype=np.float64)
        h = C.c_void_p()
        abi.check(self._lib.octlb_forest_create(abi.f64p(dom), int(bricks[0]), int(bricks[1]),
                                                int(bricks[2]), C.byref(h)), "forest create")
        self._h = h
        self.domain = dom.copy()
        self.bricks = tuple(int(b) for b in bricks)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._lib.octlb_forest_destroy(h)
            self._h = None

    @property
    def num_octants(self) -> int:
        return int(self._lib.octlb_forest_num_octants(self._h))

    def refine_uniform(self, level: int):
        big = np.array([-1e300, -1e300, -1e300, 1e300, 1e300, 1e300])
        return self.refine_box(big, level)

    def refine_box(self, box, max_level: int):
        b = np.ascontiguousarray(box, dtype=np.float64)
        abi.check(self._lib.octlb_forest_refine_box(self._h, abi.f64p(b), int(max_level)), "refine")
        return self

    def refine_marked(self, mark):
        m = np.ascontiguousarray(mark, dtype=np.uint8)
        abi.check(self._lib.octlb_forest_refine_marked(self._h, abi.u8p(m), int(m.size)), "refine")
        return self

    def balance(self):
        abi.check(self._lib.octlb_forest_balance(self._h), "balance")
        return self

    def topology(self, n: int = 8, halo: int = 1) -> abi.TopologyArrays:
        t = abi.Topology()
        abi.check(self._lib.octlb_forest_topology(self._h, int(n), int(halo), C.byref(t)), "topology")
        nb = t.num_blocks
        level = np.ctypeslib.as_array(t.level, shape=(nb,)).copy()
        bounds = np.ctypeslib.as_array(t.bounds, shape=(nb * 6,)).copy()
        same = (np.ctypeslib.as_array(t.same_level, shape=(t.num_same_level * 3,)).copy()
                if t.num_same_level else np.zeros(0, np.int32))
        tb = (np.ctypeslib.as_array(t.tree_boundary, shape=(t.num_tree_boundary * 2,)).copy()
              if t.num_tree_boundary else np.zeros(0, np.int32))
        cfs = []
        for i in range(t.num_coarse_fine):
            src = t.coarse_fine[i]
            cf = abi.CfFace()
            C.memmove(C.byref(cf), C.byref(src), C.sizeof(abi.CfFace))
            cfs.append(cf)
        return abi.TopologyArrays(n, halo, level, bounds, same, cfs, tb)


def single_block_topology(n: int, halo: int, edge: float = 1.0) -> abi.TopologyArrays:
    """One octant covering [0,edge]^3: every face is a tree-boundary face."""
    bounds = np.array([[0.0, 0.0, 0.0, edge, edge, edge]])
    tb = np.array([[0, d] for d in range(6)], dtype=np.int32)
    return abi.TopologyArrays(n, halo, np.zeros(1, np.int32), bounds, np.zeros((0, 3), np.int32),
                              [], tb)
