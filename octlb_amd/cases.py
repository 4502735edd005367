"""Case assembly: topology, BC specs, BcKind stamping, Bouzidi links, init.

Host-side setup that the reference does in its case constructors
(examples/cavity3d/cavity3d_case.h:139-194,
examples/cylinder3d/cylinder3d_case.h:392-456) and installers
(bc_installer.h:72-138, lattice_material_init.cpp:11-33,
bouzidi_link_data.cpp:139-233).  Geometry is analytic (exact circle, Gaussian
blobs) instead of CGAL voxelisation.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import abi
from .forest import Forest, single_block_topology

C = np.array([[0, 0, 0], [-1, 0, 0], [0, -1, 0], [0, 0, -1], [-1, -1, 0], [-1, 1, 0], [-1, 0, -1],
              [-1, 0, 1], [0, -1, -1], [0, -1, 1], [1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 0],
              [1, -1, 0], [1, 0, 1], [1, 0, -1], [0, 1, 1], [0, 1, -1]], dtype=np.int64)
OPP = np.array([0, 10, 11, 12, 13, 14, 15, 16, 17, 18, 1, 2, 3, 4, 5, 6, 7, 8, 9])

MAT_FLUID, MAT_SOLID, MAT_BOUNDARY = 0, 1, 2  # MaterialKind (geometry_types.h:17-21)


@dataclass
class Case:
    name: str
    topo: abi.TopologyArrays
    omega: float
    handler: int
    const_rho: int
    bouzidi: int
    snapshot: int
    specs: list
    kinds: np.ndarray                      # uint8 [nb, n, n, n]
    link: np.ndarray = field(default_factory=lambda: np.zeros((0, 3), np.int32))
    q: np.ndarray = field(default_factory=lambda: np.zeros(0))
    init_rho0: float = 1.0
    init_u0: tuple = (0.0, 0.0, 0.0)
    init_mask: int = 1 << abi.KIND_BULK
    init_field: tuple | None = None        # (rho [nb*n^3], u [nb*n^3*3])
    drag_bbox: np.ndarray | None = None
    area: float = 1.0
    u_ref: float = 1.0
    stats_scope: int = 0
    info: dict = field(default_factory=dict)

    def params(self, precision: int = 64) -> abi.Params:
        return abi.make_params(self.omega, precision=precision, handler=self.handler,
                               const_rho=self.const_rho, stats_scope=self.stats_scope,
                               bouzidi=self.bouzidi, snapshot=self.snapshot, specs=self.specs)

    def apply(self, sim):
        """Stamp this case onto a solver handle: cell kinds, Bouzidi links
        and the initial populations (the case constructors' install order,
        cylinder3d_case.h:392-456).  Returns ``sim``."""
        sim.set_bc_kinds(self.kinds)
        if self.link.shape[0] > 0:
            sim.set_links(self.link, self.q)
        if self.init_field is not None:
            sim.init_field(*self.init_field)
        else:
            sim.init_equilibrium(self.init_rho0, self.init_u0, self.init_mask)
        return sim

    @property
    def n(self):
        return self.topo.n

    @property
    def num_blocks(self):
        return self.topo.num_blocks


# ------------------------------------------------------------- installers
def bc_kind_from_spec_type(t: int) -> int:
    """BcKindFromSpecType (bc_installer.h:44-58)."""
    return {abi.BC_NO_SLIP: abi.KIND_BOUNCE_BACK, abi.BC_MOVING_LID: abi.KIND_MOVING_BOUNCE_BACK,
            abi.BC_INTERP_VELOCITY: abi.KIND_VELOCITY_DIRICHLET, abi.BC_OUTFLOW: abi.KIND_OUTFLOW,
            abi.BC_INTERP_PRESSURE: abi.KIND_PRESSURE_DIRICHLET}[t]


def find_spec_type(specs, face: int) -> int:
    for s in specs:
        if s.face == face:
            return s.type
    return abi.BC_NO_SLIP  # default DomainBcSpec{}


def stamp_tree_boundary(kinds: np.ndarray, tree_boundary: np.ndarray, specs) -> None:
    """StampTreeBoundaryCells (bc_installer.h:96-138): face slabs take the
    spec's kind in list order, skipping kSolid cells."""
    for b, d in tree_boundary:
        kind = bc_kind_from_spec_type(find_spec_type(specs, int(d)))
        kb = kinds[b]
        axis, side = int(d) // 2, int(d) & 1
        idx = [slice(None)] * 3
        idx[axis] = -1 if side else 0
        slab = kb[tuple(idx)]
        slab[slab != abi.KIND_SOLID] = kind


def material_to_kinds(material: np.ndarray) -> np.ndarray:
    """BcKindFromMaterial (bc_installer.h:31-41)."""
    lut = np.array([abi.KIND_BULK, abi.KIND_SOLID, abi.KIND_BOUZIDI], dtype=np.uint8)
    return lut[material]


# ------------------------------------------------------------- unit converter
def unit_converter(resolution, tau, length, velocity, nu, rho=1.0):
    """UnitConverter (unit_converter.cpp:8-26): returns (omega, u_lattice)."""
    inv_cs2 = 3.0
    dx = length / float(resolution)
    dt = (tau - 0.5) / inv_cs2 * dx * dx / nu
    conv_u = dx / dt
    return 1.0 / tau, velocity / conv_u


# ------------------------------------------------------------- cavity (C1)
def cavity(n_lat: int = 64, tau: float = 0.6, u_lid: float = 0.05, halo: int = 3, nu: float = None) -> Case:
    """Lid-driven cavity in the reference's single-block form
    (cavity3d_case.h:139-194): one n_lat^3 block with OpenLB padding 3,
    ConstRhoBGK bulk, interpolated-velocity (Skordos FD) on all six faces,
    lid = u on the y-max face, padded-block rotate streaming.  ``nu`` given:
    UnitConverter(n_lat - 1, tau, 1, 1, nu) exactly (the lattice lid speed
    follows from it, e.g. OpenLbCavity3dDefaults = (30, 0.509, 1, 1, 0.001),
    unit_converter.cpp:28-31); otherwise nu is chosen so the lid is u_lid."""
    res = n_lat - 1
    if nu is None:
        nu = (tau - 0.5) / 3.0 * (1.0 / res) / u_lid  # gives u_lattice = u_lid
    omega, u_lat = unit_converter(res, tau, 1.0, 1.0, nu)
    specs = [abi.make_spec(abi.XMIN, abi.BC_INTERP_VELOCITY),
             abi.make_spec(abi.XMAX, abi.BC_INTERP_VELOCITY),
             abi.make_spec(abi.YMIN, abi.BC_INTERP_VELOCITY),
             abi.make_spec(abi.ZMIN, abi.BC_INTERP_VELOCITY),
             abi.make_spec(abi.ZMAX, abi.BC_INTERP_VELOCITY),
             abi.make_spec(abi.YMAX, abi.BC_INTERP_VELOCITY, u_wall=(u_lat, 0.0, 0.0))]
    topo = single_block_topology(n_lat, halo)
    kinds = np.zeros((1, n_lat, n_lat, n_lat), np.uint8)
    stamp_tree_boundary(kinds, topo.tree_boundary, specs)
    return Case(name=f"cavity{n_lat}", topo=topo, omega=omega, handler=abi.HANDLER_CONCRETE,
                const_rho=1, bouzidi=0, snapshot=0, specs=specs, kinds=kinds,
                init_mask=0xFF, u_ref=u_lat, info={"u_lid": u_lat, "tau": tau, "resolution": res})


# ------------------------------------------------------------- cylinder (C2)
def _ray_circle(px, py, dx, dy, xc, yc, r):
    """Smallest s in (0,1] with |p + s d - c| = r, or nan."""
    fx, fy = px - xc, py - yc
    a = dx * dx + dy * dy
    b = 2.0 * (fx * dx + fy * dy)
    c = fx * fx + fy * fy - r * r
    with np.errstate(invalid="ignore", divide="ignore"):
        disc = b * b - 4.0 * a * c
        sq = np.sqrt(np.maximum(disc, 0.0))
        s1 = (-b - sq) / (2.0 * a)
        s2 = (-b + sq) / (2.0 * a)
    s = np.where((s1 > 0.0) & (s1 <= 1.0), s1, np.where((s2 > 0.0) & (s2 <= 1.0), s2, np.nan))
    return np.where((a > 0.0) & (disc >= 0.0), s, np.nan)


def cell_centres(topo: abi.TopologyArrays):
    n = topo.n
    b = topo.bounds
    cw = (b[:, 3] - b[:, 0]) / n
    idx = (np.arange(n) + 0.5)
    x = b[:, 0, None] + cw[:, None] * idx[None, :]
    y = b[:, 1, None] + cw[:, None] * idx[None, :]
    z = b[:, 2, None] + cw[:, None] * idx[None, :]
    X = np.broadcast_to(x[:, :, None, None], (len(b), n, n, n))
    Y = np.broadcast_to(y[:, None, :, None], (len(b), n, n, n))
    Z = np.broadcast_to(z[:, None, None, :], (len(b), n, n, n))
    return X, Y, Z, cw


def cylinder_channel(bricks_x: int = 4, base_level: int = 2, fine: bool = True, n: int = 8,
                     bricks_yz: int = 1,
                     xc: float = 1.0, radius: float = 0.125, u_in: float = 0.04, tau: float = 0.55,
                     refine_lo: float = None, refine_hi: float = None, ramp_end_t: float = 0.0,
                     refine_y: tuple = (-1.0, 2.0), open_ends: bool = True) -> Case:
    """Channel with a z-spanning cylinder on a 2-level brick forest.

    Coarse blocks at ``base_level`` fill [0,bricks_x]x[0,1]x[0,1]; blocks
    overlapping x in [xc-2D, xc+6D] (and y in ``refine_y``) are refined once.  Inlet x-min =
    interpolated velocity (uniform u_in), outlet x-max = interpolated pressure
    (rho 1), lateral faces have no spec (default no-slip -> kBounceBack,
    bc_installer.h:44-58).  Cylinder cells are solid, its surface cells
    kBouzidi; q is the exact ray-circle fraction of every cut fluid link.
    A single relaxation rate is used on all levels (time_loop.h:183-187).

    ``open_ends=False`` drops the inlet and outlet: every face is a
    bounce-back wall and the fluid starts at feq(1, (u_in, 0, 0)) -- the
    closed box the discrete adjoint differentiates (no FD boundary cells)."""
    dom = np.array([0.0, 0.0, 0.0, float(bricks_x) / bricks_yz, 1.0, 1.0])
    f = Forest(dom, (bricks_x, bricks_yz, bricks_yz))
    f.refine_uniform(base_level)
    d = 2.0 * radius
    lo = xc - 2.0 * d if refine_lo is None else refine_lo
    hi = xc + 6.0 * d if refine_hi is None else refine_hi
    if fine:
        eps = 1e-9
        f.refine_box([lo + eps, refine_y[0] + eps, -1.0, hi - eps, refine_y[1] - eps, 2.0], base_level + 1)
        f.balance()
    topo = f.topology(n, 1)
    X, Y, Z, cw = cell_centres(topo)
    inside = (X - xc) ** 2 + (Y - 0.5) ** 2 < radius * radius
    nb = topo.num_blocks
    # material: solid inside; solid with a fluid D3Q19 neighbour -> boundary
    material = np.where(inside, MAT_SOLID, MAT_FLUID).astype(np.int64)
    surf = np.zeros_like(inside)
    for i in range(1, 19):
        px = X + C[i, 0] * cw[:, None, None, None]
        py = Y + C[i, 1] * cw[:, None, None, None]
        nin = (px - xc) ** 2 + (py - 0.5) ** 2 < radius * radius
        surf |= inside & ~nin
    material[surf] = MAT_BOUNDARY
    kinds = material_to_kinds(material)
    specs = [abi.make_spec(abi.XMIN, abi.BC_INTERP_VELOCITY, inlet=abi.INLET_UNIFORM,
                           inlet_u=(u_in, 0.0, 0.0), ramp_end_t=ramp_end_t),
             abi.make_spec(abi.XMAX, abi.BC_INTERP_PRESSURE, rho_target=1.0)] if open_ends else []
    stamp_tree_boundary(kinds, topo.tree_boundary, specs)
    # Bouzidi links of material-fluid cells whose in-block pull source x - c_i
    # is solid-like (bouzidi_link_data.cpp:188-225), q in (1e-8, 1-1e-8)
    links, qs = [], []
    fluid = material == MAT_FLUID
    for i in range(1, 19):
        sx = np.arange(n)[:, None, None] - C[i, 0]
        sy = np.arange(n)[None, :, None] - C[i, 1]
        sz = np.arange(n)[None, None, :] - C[i, 2]
        inb = (sx >= 0) & (sx < n) & (sy >= 0) & (sy < n) & (sz >= 0) & (sz < n)
        sxc, syc, szc = np.clip(sx, 0, n - 1), np.clip(sy, 0, n - 1), np.clip(sz, 0, n - 1)
        src_solid = material[:, sxc, syc, szc] != MAT_FLUID
        cand = fluid & inb[None] & src_solid
        if not cand.any():
            continue
        bb, ix, iy, iz = np.nonzero(cand)
        s = _ray_circle(X[bb, ix, iy, iz], Y[bb, ix, iy, iz], -C[i, 0] * cw[bb], -C[i, 1] * cw[bb],
                        xc, 0.5, radius)
        ok = np.isfinite(s) & (s > 1e-8) & (s < 1.0 - 1e-8)
        cell = (ix * n + iy) * n + iz
        links.append(np.stack([bb[ok], cell[ok], np.full(ok.sum(), i)], axis=1))
        qs.append(s[ok])
    link = np.concatenate(links).astype(np.int32) if links else np.zeros((0, 3), np.int32)
    q = np.concatenate(qs) if qs else np.zeros(0)
    fine_dx = float(cw.min())
    pad = 3.0 * fine_dx
    bbox = np.array([xc - radius - pad, 0.5 - radius - pad, 0.0 - pad,
                     xc + radius + pad, 0.5 + radius + pad, 1.0 + pad])
    area = (d / fine_dx) * (1.0 / fine_dx)  # (D/dx)(z_span/dx), test_cylinder3d_amr.cpp:99-105
    return Case(name="cylinder", topo=topo, omega=1.0 / tau, handler=abi.HANDLER_CONCRETE,
                const_rho=0, bouzidi=1, snapshot=1, specs=specs, kinds=kinds, link=link, q=q,
                init_mask=1 << abi.KIND_BULK, drag_bbox=bbox, area=area, u_ref=u_in,
                init_u0=(0.0, 0.0, 0.0) if open_ends else (u_in, 0.0, 0.0),
                info={"fine_dx": fine_dx, "D": d, "material_fluid": fluid})


# ------------------------------------------------------ synthetic AMR (C3)
def gaussian_blobs(seed: int, count: int = 8):
    """Centres U[0,1]^3 and widths sigma U[0.05, 0.15] of the blob field."""
    rng = np.random.default_rng(seed)
    return rng.random((count, 3)), rng.uniform(0.05, 0.15, count)


def blob_field(points, centres, sigmas):
    """phi(x) = sum_b exp(-|x - c_b|^2 / (2 sigma_b^2))."""
    d2 = ((points[:, None, :] - centres[None, :, :]) ** 2).sum(-1)
    return np.exp(-d2 / (2.0 * sigmas[None, :] ** 2)).sum(-1)


def fourier_spectrum(seed: int, kmax: int = 4):
    """Random complex amplitudes a_k on the integer wave vectors 0 < |k| <=
    kmax, each projected onto the plane normal to k (Helmholtz projection:
    every mode, hence the field, is divergence-free).  Returned dense on the
    (2 kmax + 1)^3 lattice of k, zero outside the ball."""
    rng = np.random.default_rng(seed)
    r = np.arange(-kmax, kmax + 1)
    KX, KY, KZ = np.meshgrid(r, r, r, indexing="ij")
    K = np.stack([KX, KY, KZ], -1).astype(np.float64)
    k2 = (K ** 2).sum(-1)
    A = rng.normal(size=K.shape) + 1j * rng.normal(size=K.shape)
    kk = np.where(k2[..., None] > 0, K / np.sqrt(np.maximum(k2, 1))[..., None], 0.0)
    A = A - (A * kk).sum(-1, keepdims=True) * kk
    A[(k2 == 0) | (k2 > kmax * kmax)] = 0.0
    return A


def fourier_velocity(topo: abi.TopologyArrays, A, kmax: int = 4):
    """u(x) = Re sum_k a_k exp(2 pi i k.x) at every cell centre, evaluated by
    contracting one axis at a time over each block's 8 x 8 x 8 centres."""
    X, Y, Z, _ = cell_centres(topo)
    r = np.arange(-kmax, kmax + 1)
    xs, ys, zs = X[:, :, 0, 0], Y[:, 0, :, 0], Z[:, 0, 0, :]
    Ex = np.exp(2j * np.pi * xs[..., None] * r)  # (nb, n, 2kmax+1)
    Ey = np.exp(2j * np.pi * ys[..., None] * r)
    Ez = np.exp(2j * np.pi * zs[..., None] * r)
    t1 = np.einsum("bia,aecd->biecd", Ex, A, optimize=True)
    t2 = np.einsum("bje,biecd->bijcd", Ey, t1, optimize=True)
    u = np.einsum("bkc,bijcd->bijkd", Ez, t2, optimize=True).real
    return u


def synthetic_forest(bricks: int = 8, max_level: int = 2, frac: tuple = (0.5, 0.25), seed: int = 7) -> Forest:
    """The forest of synthetic_amr: unit cube of bricks^3 level-0 blocks
    refined where a seeded sum of 8 Gaussian blobs is largest."""
    centres, sigmas = gaussian_blobs(seed)
    f = Forest([0.0, 0.0, 0.0, 1.0, 1.0, 1.0], (bricks, bricks, bricks))
    for lv in range(max_level):
        t = f.topology(8, 1)
        b = t.bounds
        mid = 0.5 * (b[:, :3] + b[:, 3:])
        phi = blob_field(mid, centres, sigmas)
        vol = np.prod(b[:, 3:] - b[:, :3], axis=1)
        at = t.level == lv
        order = np.argsort(-phi[at])
        target = frac[lv] if lv < len(frac) else frac[-1]
        cum = np.cumsum(vol[at][order])
        take = order[: int(np.searchsorted(cum, target)) + 1] if cum.size else order[:0]
        mark = np.zeros(t.num_blocks, np.uint8)
        mark[np.flatnonzero(at)[take]] = 1
        f.refine_marked(mark)
        f.balance()
    return f


def walls_kinds(topo) -> np.ndarray:
    """BcKind bytes of a forest with no-slip walls on all six faces."""
    kinds = np.zeros((topo.num_blocks, topo.n, topo.n, topo.n), np.uint8)
    stamp_tree_boundary(kinds, topo.tree_boundary, [])
    return kinds


def synthetic_amr(bricks: int = 8, max_level: int = 2, frac: tuple = (0.5, 0.25), seed: int = 7,
                  amp: float = 0.01, kmax: int = 4, tau: float = 0.6) -> Case:
    """Unit cube of bricks^3 level-0 blocks refined where a seeded sum of 8
    Gaussian blobs is largest: level-0 octants whose centre value lies in the
    top frac[0] of the domain volume are refined, then level-1 octants
    likewise until frac[1] of the volume is at level 2; 2:1 balanced.  The
    initial velocity is a divergence-free random Fourier field (|k| <= kmax)
    scaled to max |u| = amp, loaded as feq(1, u).  Walls on all six faces
    (no spec: no-slip -> kBounceBack, bc_installer.h:44-58)."""
    topo = synthetic_forest(bricks, max_level, frac, seed).topology(8, 1)
    kinds = walls_kinds(topo)
    u = fourier_velocity(topo, fourier_spectrum(seed + 1, kmax), kmax)
    u *= amp / np.sqrt((u ** 2).sum(-1)).max()
    rho = np.ones(kinds.shape)
    return Case(name="synthetic_amr", topo=topo, omega=1.0 / tau, handler=abi.HANDLER_CONCRETE,
                const_rho=0, bouzidi=0, snapshot=0, specs=[], kinds=kinds,
                init_field=(rho.reshape(-1), u.reshape(-1)),
                info={"seed": seed, "bricks": bricks, "max_level": max_level, "frac": frac})


# ------------------------------------------------------ cavity ensemble (C4)
def cavity_blocked(cells: int, tau: float = 0.6, u_lid: float = 0.05, const_rho: int = 1,
                   x0: float = 0.0) -> Case:
    """Lid-driven cavity of cells^3 cells as a uniform forest of 8^3 blocks
    (halo 1): interpolated-velocity (Skordos FD) walls on all six faces, lid
    u_lid (lattice units) along x on the y-max face, ConstRhoBGK bulk when
    const_rho, start at rest.  The domain is [x0, x0+1] x [0,1]^2, so members
    of an ensemble can sit side by side."""
    if cells % 8:
        raise ValueError("cells must be a multiple of 8")
    k = cells // 8
    f = Forest([x0, 0.0, 0.0, x0 + 1.0, 1.0, 1.0], (k, k, k))
    topo = f.topology(8, 1)
    specs = [abi.make_spec(face, abi.BC_INTERP_VELOCITY) for face in
             (abi.XMIN, abi.XMAX, abi.YMIN, abi.ZMIN, abi.ZMAX)]
    specs.append(abi.make_spec(abi.YMAX, abi.BC_INTERP_VELOCITY, u_wall=(u_lid, 0.0, 0.0)))
    kinds = np.zeros((topo.num_blocks, 8, 8, 8), np.uint8)
    stamp_tree_boundary(kinds, topo.tree_boundary, specs)
    return Case(name=f"cavity{cells}", topo=topo, omega=1.0 / tau, handler=abi.HANDLER_CONCRETE,
                const_rho=const_rho, bouzidi=0, snapshot=0, specs=specs, kinds=kinds,
                init_mask=0xFF, u_ref=u_lid, info={"cells": cells, "tau": tau, "u_lid": u_lid})


@dataclass
class Ensemble:
    topo: abi.TopologyArrays
    member_begin: np.ndarray   # int32 [members + 1]
    members: list              # the member Cases, in block order
    kinds: np.ndarray
    fast_stats: int = 0

    def params_for(self, precision: int = 64):
        out = []
        for c in self.members:
            p = c.params(precision)
            p.fast_stats = self.fast_stats
            out.append(p)
        return out


def ensemble(members, fast_stats: int = 0) -> Ensemble:
    """Concatenate single-level member cases into one block list (block ids
    of member m offset by the blocks before it; no face joins members)."""
    n, halo = members[0].topo.n, members[0].topo.halo
    level, bounds, same, tb, begin = [], [], [], [], [0]
    for c in members:
        t = c.topo
        if t.n != n or t.halo != halo or t.num_cf or t.max_level:
            raise ValueError("ensemble members must be single-level forests of the same block shape")
        off = begin[-1]
        level.append(t.level)
        bounds.append(t.bounds)
        s = t.same_level.copy()
        s[:, 0] += off
        s[:, 2] += off
        same.append(s)
        b = t.tree_boundary.copy()
        b[:, 0] += off
        tb.append(b)
        begin.append(off + t.num_blocks)
    topo = abi.TopologyArrays(n, halo, np.concatenate(level), np.concatenate(bounds),
                              np.concatenate(same), [], np.concatenate(tb))
    kinds = np.concatenate([c.kinds for c in members])
    return Ensemble(topo=topo, member_begin=np.array(begin, np.int32), members=list(members), kinds=kinds,
                    fast_stats=fast_stats)


def cavity_ensemble(count: int = 64, sizes=(32, 48, 64, 96), tau=(0.52, 0.8), lid=(0.02, 0.08), seed: int = 11,
                    const_rho: int = 1, fast_stats: int = 1) -> Ensemble:
    """`count` cavities with cells drawn from `sizes`, tau and lid speed
    uniform in the given ranges, one seed; member m sits at x0 = 2m."""
    rng = np.random.default_rng(seed)
    ns = rng.choice(np.asarray(sizes), size=count)
    taus = rng.uniform(tau[0], tau[1], size=count)
    lids = rng.uniform(lid[0], lid[1], size=count)
    members = [cavity_blocked(int(ns[m]), float(taus[m]), float(lids[m]), const_rho, x0=2.0 * m)
               for m in range(count)]
    return ensemble(members, fast_stats=fast_stats)
