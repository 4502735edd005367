"""Triangle soups, material fields and the GPU Bouzidi link build.

``build_links`` calls ``octlb_build_links`` (the device form of
BouzidiLinkData::Build, bouzidi_link_data.cpp:139-233).  The soup helpers
produce the kinds of geometry the reference's parts are made of (closed
boxes, spheres, z-spanning cylinders), one triangle = 9 doubles (v0, v1, v2);
``material_field`` voxelises an analytic inside-test the way the reference
marks cells (solid inside, boundary = solid with a fluid D3Q19 neighbour),
in MaterialField order [block][k][j][i].
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from octlb_amd import abi

MAT_FLUID, MAT_SOLID, MAT_BOUNDARY = 0, 1, 2


def box_soup(lo, hi) -> np.ndarray:
    """Closed axis-aligned box, two triangles per face."""
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    v = np.array([[x0, y0, z0], [x1, y0, z0], [x1, y1, z0], [x0, y1, z0],
                  [x0, y0, z1], [x1, y0, z1], [x1, y1, z1], [x0, y1, z1]], np.float64)
    quads = [(0, 3, 2, 1), (4, 5, 6, 7), (0, 1, 5, 4), (3, 7, 6, 2), (0, 4, 7, 3), (1, 2, 6, 5)]
    tris = []
    for a, b, c, d in quads:
        tris.append(np.concatenate([v[a], v[b], v[c]]))
        tris.append(np.concatenate([v[a], v[c], v[d]]))
    return np.array(tris)


def sphere_soup(centre, r, segments) -> np.ndarray:
    """Latitude/longitude sphere, 2 * segments^2 triangles."""
    cx, cy, cz = centre
    th = np.linspace(0.0, np.pi, segments + 1)
    ph = np.linspace(0.0, 2.0 * np.pi, segments + 1)

    def pt(t, p):
        return np.array([cx + r * np.sin(t) * np.cos(p), cy + r * np.sin(t) * np.sin(p), cz + r * np.cos(t)])

    tris = []
    for a in range(segments):
        for b in range(segments):
            p00, p01 = pt(th[a], ph[b]), pt(th[a], ph[b + 1])
            p10, p11 = pt(th[a + 1], ph[b]), pt(th[a + 1], ph[b + 1])
            tris.append(np.concatenate([p00, p10, p01]))
            tris.append(np.concatenate([p01, p10, p11]))
    return np.array(tris)


def cylinder_soup(xc, yc, r, z0, z1, segments, rings=1) -> np.ndarray:
    """Side of a z-axis cylinder (no caps), `rings` bands along z."""
    ph = np.linspace(0.0, 2.0 * np.pi, segments + 1)
    zs = np.linspace(z0, z1, rings + 1)
    tris = []
    for a in range(rings):
        for b in range(segments):
            p = [np.array([xc + r * np.cos(ph[b + u]), yc + r * np.sin(ph[b + u]), zs[a + w]])
                 for w in (0, 1) for u in (0, 1)]
            tris.append(np.concatenate([p[0], p[1], p[3]]))
            tris.append(np.concatenate([p[0], p[3], p[2]]))
    return np.array(tris)


def material_field(topo: abi.TopologyArrays, inside) -> np.ndarray:
    """MaterialKind per cell in MaterialField order [block][k][j][i]:
    inside(x, y, z) -> solid; solid cells with a fluid D3Q19 neighbour
    inside the block become boundary."""
    n = topo.n
    b = topo.bounds
    nb = b.shape[0]
    idx = (np.arange(n) + 0.5) / n
    w = b[:, 3:] - b[:, :3]
    X = b[:, 0, None, None, None] + w[:, 0, None, None, None] * idx[None, None, None, :]
    Y = b[:, 1, None, None, None] + w[:, 1, None, None, None] * idx[None, None, :, None]
    Z = b[:, 2, None, None, None] + w[:, 2, None, None, None] * idx[None, :, None, None]
    X, Y, Z = np.broadcast_arrays(X, Y, Z)  # [block][k][j][i]
    solid = np.asarray(inside(X, Y, Z), bool)
    mat = np.where(solid, MAT_SOLID, MAT_FLUID).astype(np.uint8)
    from octlb_amd.cases import C as CV
    pad = np.pad(solid, ((0, 0), (1, 1), (1, 1), (1, 1)), constant_values=True)
    fluid_nb = np.zeros_like(solid)
    for i in range(1, 19):
        ex, ey, ez = CV[i]
        fluid_nb |= ~pad[:, 1 + ez:1 + ez + n, 1 + ey:1 + ey + n, 1 + ex:1 + ex + n]
    mat[solid & fluid_nb] = MAT_BOUNDARY
    assert mat.shape == (nb, n, n, n)
    return mat


def build_links(bounds, n: int, material, triangles):
    """Device link build; returns (links int32 [L, 3] = block, cell
    (ix*n+iy)*n+iz, population; q float64 [L])."""
    h = abi.lib()
    bounds = np.ascontiguousarray(bounds, np.float64).reshape(-1, 6)
    material = np.ascontiguousarray(material, np.uint8)
    tri = np.ascontiguousarray(triangles, np.float64).reshape(-1, 9)
    nb = bounds.shape[0]
    if material.size != nb * n ** 3:
        raise ValueError("material must hold n^3 bytes per block")
    cnt = C.c_int64(0)
    args = (nb, n, abi.f64p(bounds), abi.u8p(material.reshape(-1)), tri.shape[0],
            abi.f64p(tri) if tri.size else None)
    abi.check(h.octlb_build_links(*args, 0, C.byref(cnt), None, None), "octlb_build_links (count)")
    L = cnt.value
    link = np.zeros((max(L, 1), 3), np.int32)
    q = np.zeros(max(L, 1))
    if L:
        abi.check(h.octlb_build_links(*args, L, C.byref(cnt), abi.i32p(link), abi.f64p(q)), "octlb_build_links")
    return link[:L], q[:L]
