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


PART_INTERNAL_CHANNEL, PART_EXTERNAL_OBSTACLE = 0, 1  # GeometryPartRole


def reference_box_soup(lo, hi, inward=False) -> np.ndarray:
    """The reference's test box (AddAxisAlignedBox, tests/unit/mesh/
    geometry_fixtures.h:39-70): its twelve triangles in its order and
    orientation, as they are -- they do not close the box (one x-max
    triangle appears twice in both orientations, one triangle cuts through
    the inside, half of the x-max and z-max faces is missing), and the
    reference's ray-parity inside test classifies against exactly this soup."""
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    p = {"000": (x0, y0, z0), "100": (x1, y0, z0), "010": (x0, y1, z0), "110": (x1, y1, z0),
         "001": (x0, y0, z1), "101": (x1, y0, z1), "011": (x0, y1, z1), "111": (x1, y1, z1)}
    order = [("000", "100", "010"), ("100", "110", "010"), ("100", "110", "101"), ("100", "101", "110"),
             ("110", "011", "101"), ("110", "011", "111"), ("010", "110", "011"), ("010", "011", "001"),
             ("001", "011", "101"), ("001", "101", "100"), ("000", "010", "001"), ("000", "001", "100")]
    tris = [p[a] + (p[c] + p[b] if inward else p[b] + p[c]) for a, b, c in order]
    return np.array(tris, np.float64)


def reference_sphere_soup(cx, cy, cz, r, segments) -> np.ndarray:
    """CenteredSphere (tests/unit/mesh/geometry_fixtures.h:88-116): the
    same angles and operation order, so the same vertex doubles."""
    import math
    tris = []
    for lat in range(segments):
        th0 = float(lat) * math.pi / segments
        th1 = float(lat + 1) * math.pi / segments
        for lon in range(segments):
            ph0 = float(lon) * 2.0 * math.pi / segments
            ph1 = float(lon + 1) * 2.0 * math.pi / segments

            def sph(t, f):
                st = math.sin(t)
                return (cx + r * st * math.cos(f), cy + r * st * math.sin(f), cz + r * math.cos(t))

            p00, p01, p10, p11 = sph(th0, ph0), sph(th0, ph1), sph(th1, ph0), sph(th1, ph1)
            tris.append(p00 + p10 + p01)
            tris.append(p01 + p10 + p11)
    return np.array(tris, np.float64)


def voxelize_part(bounds, n: int, soup, role: int, inner=None):
    """Device voxelize_part (part_voxelizer.cpp:71-114): MaterialKind bytes
    [block][k][j][i] of one part; returns (material, cells that needed the
    exact-arithmetic fallback)."""
    h = abi.lib()
    bounds = np.ascontiguousarray(bounds, np.float64).reshape(-1, 6)
    tri = np.ascontiguousarray(soup, np.float64).reshape(-1, 9)
    inn = np.ascontiguousarray(inner if inner is not None else np.zeros((0, 9)), np.float64).reshape(-1, 9)
    nb = bounds.shape[0]
    out = np.zeros((nb, n, n, n), np.uint8)
    nexact = C.c_int64(0)
    abi.check(h.octlb_voxelize_part(nb, n, abi.f64p(bounds), tri.shape[0], abi.f64p(tri) if tri.size else None,
                                    inn.shape[0], abi.f64p(inn) if inn.size else None, int(role),
                                    abi.u8p(out.reshape(-1)), C.byref(nexact)), "octlb_voxelize_part")
    return out, int(nexact.value)


def merge_material(base, overlay, role: int) -> None:
    """merge_material_field (part_voxelizer.cpp:116-146), in place: an
    internal channel overwrites every cell, an external obstacle only the
    cells it classifies solid or boundary."""
    if base.shape != overlay.shape:
        raise ValueError("MaterialField merge size mismatch")
    if role == PART_EXTERNAL_OBSTACLE:
        claim = overlay != MAT_FLUID
        base[claim] = overlay[claim]
    else:
        base[...] = overlay


def build_material(bounds, n: int, parts, voxelize=None):
    """GeometryEngine::build's voxelize-and-merge (geometry_engine.cpp:28-39,
    without the forest refinement): parts are dicts with soup, role,
    priority and (channels) inner, merged in ascending priority, equal
    priorities in the given order.  voxelize(bounds, n, soup, role, inner)
    -> material defaults to the device voxelizer."""
    if not parts:
        raise ValueError("GeometryAssembly has no parts")
    if voxelize is None:
        voxelize = lambda b, m, s, r, i: voxelize_part(b, m, s, r, i)[0]  # noqa: E731
    nb = np.asarray(bounds).reshape(-1, 6).shape[0]
    merged = np.zeros((nb, n, n, n), np.uint8)
    for part in sorted(parts, key=lambda q: q.get("priority", 0)):
        field = voxelize(bounds, n, part["soup"], part["role"], part.get("inner"))
        merge_material(merged, field, part["role"])
    return merged


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
