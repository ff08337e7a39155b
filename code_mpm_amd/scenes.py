"""Seeded scene generators: the world set-up that the reference writes as
``PhysicsEngine::initializeWorld_*`` functions (e.g.
PhysicsEngine_initializeWorld_Classic_Bar.cpp:4-256,
PhysicsEngine_initializeWorld_Classic_Cellular_Langrand_Hexagonal.cpp:4-301).

A Scene holds plain numpy arrays and can be applied to any backend that has
the Engine set-up methods.  Conventions kept from the reference:

- grid nodes at (i * L) / cells, fixed masks chosen by node position with the
  reference idiom ``fabs(x - x0) < 1.5 h`` where a scene says so;
- particles at voxel centres of a sub-cell lattice (Canvas2D_CC.cpp:26:
  ``(i + 0.5) * offset``), volume offset^3 and mass density * volume
  (IW_Bar:154-157);
- the top platen is displacement-controlled particles with multiplier -1
  (IW_Cellular_Langrand_Hexagonal:213-219); the bottom support is fixed-y
  nodes whose reaction is the measured force (DP:125-131).

Particles are emitted cell by cell (x fastest over cells, then the 2x2x2
sub-positions), so the 8 particles of a cell are adjacent in memory.
"""
from dataclasses import dataclass, field

import numpy as np

from .capi import ELASTIC, PLASTIC, VONMISESHARDENING

AXIS_X, AXIS_Y, AXIS_Z = 1, 2, 4


@dataclass
class Scene:
    name: str
    length: np.ndarray
    cells: np.ndarray
    fixed: np.ndarray
    friction: np.ndarray
    materials: list
    fields: dict
    material: np.ndarray
    dt: float
    time_end: float
    dc_index: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int64))
    dc_axes: np.ndarray = field(default_factory=lambda: np.zeros(0, np.uint8))
    dc_mult: np.ndarray = field(default_factory=lambda: np.zeros(0, np.float32))
    monitors: dict = field(default_factory=dict)
    timeline_t: np.ndarray = field(default_factory=lambda: np.zeros(0))
    timeline_v: np.ndarray = field(default_factory=lambda: np.zeros((0, 3)))
    damping: float = 0.0
    mass_min: float = 0.0
    time0: float = 0.0
    console_interval: float = None
    info: dict = field(default_factory=dict)

    @property
    def n(self):
        return int(self.material.shape[0])

    @property
    def nnodes(self):
        c = np.asarray(self.cells) + 1
        return int(c[0] * c[1] * c[2])

    @property
    def h(self):
        return np.asarray(self.length, np.float64) / np.asarray(self.cells, np.float64)

    def apply(self, be):
        be.set_grid(self.length, self.cells)
        be.set_node_boundary(self.fixed, self.friction)
        be.set_materials(self.materials)
        be.set_particle_count(self.n)
        for name, arr in self.fields.items():
            be.set_particle(name, arr)
        be.set_particle_material(self.material)
        be.set_displacement_control(self.dc_index, self.dc_axes, self.dc_mult)
        for kind, idx in self.monitors.items():
            be.set_monitor_list(kind, idx)
        be.set_timeline(self.timeline_t, self.timeline_v)
        be.set_time(self.time0, self.dt, self.time_end)
        be.set_damping(self.damping, self.mass_min)


def node_positions(length, cells):
    """(3, nnodes) node coordinates (i * L) / cells, index ix + iy*nx + iz*nx*ny."""
    nn = np.asarray(cells) + 1
    ax = [np.arange(nn[a], dtype=np.float64) * float(length[a]) / float(cells[a]) for a in range(3)]
    z, y, x = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
    return np.stack([x.ravel(), y.ravel(), z.ravel()])


def lattice_in_cells(cell_idx, h):
    """Positions of the 2x2x2 sub-cell particles (offsets 0.25 h and 0.75 h)
    of the given cells (k, 3) int array, emitted cell by cell -> (3, 8k)."""
    sub = np.array([[ox, oy, oz] for oz in (0.25, 0.75) for oy in (0.25, 0.75) for ox in (0.25, 0.75)])
    pos = (cell_idx[:, None, :].astype(np.float64) + sub[None, :, :]) * np.asarray(h)[None, None, :]
    return np.ascontiguousarray(pos.reshape(-1, 3).T)


def default_fields(x, v, volume, density):
    n = x.shape[1]
    F = np.zeros((9, n))
    F[0] = F[4] = F[8] = 1.0
    V0 = np.full(n, float(volume))
    return {
        "POSITION": np.ascontiguousarray(x),
        "VELOCITY": np.ascontiguousarray(v),
        "MASS": density * V0,
        "VOLUME": V0.copy(),
        "VOLUME_INITIAL": V0,
        "DEFORMATION_GRADIENT": F,
    }


def elastic_bar(cells=(64, 8, 8), h=1.5625e-3, E=200.0e9, nu=0.3, rho=7850.0, tip_speed=1.0):
    """Linear-elastic 3D cantilever bar filling a 64x8x8-cell grid
    (h = 1.5625e-3 m), 2x2x2 particles per cell (32,768), the x = 0 node
    plane fixed on all axes, the particles of the last cell column (x > 63 h)
    given v = (0, 0, -tip_speed); dt = 0.1 h / sqrt(E / rho)."""
    cells = np.asarray(cells, dtype=np.int32)
    length = cells * h
    hh = length / cells
    cz, cy, cx = np.meshgrid(np.arange(cells[2]), np.arange(cells[1]), np.arange(cells[0]), indexing="ij")
    cidx = np.stack([cx.ravel(), cy.ravel(), cz.ravel()], axis=1)
    x = lattice_in_cells(cidx, hh)
    n = x.shape[1]
    v = np.zeros((3, n))
    tip = x[0] > (cells[0] - 1) * hh[0]
    v[2, tip] = -tip_speed
    offset = hh[0] / 2.0
    fields = default_fields(x, v, offset ** 3, rho)
    nodes = node_positions(length, cells)
    fixed = np.where(nodes[0] == 0.0, AXIS_X | AXIS_Y | AXIS_Z, 0).astype(np.uint8)
    c = np.sqrt(E / rho)
    dt = 0.1 * hh[0] / c
    mats = [dict(type=ELASTIC, density=rho, elastic_modulus=E, poisson_ratio=nu)]
    root = np.nonzero(x[0] < 2.0 * hh[0])[0]
    return Scene(
        name="elastic_bar_64x8x8", length=length, cells=cells, fixed=fixed, friction=np.zeros(fixed.shape),
        materials=mats, fields=fields, material=np.zeros(n, np.int32), dt=dt, time_end=1000 * dt,
        monitors={"DISPLACEMENT": np.nonzero(tip)[0], "MOMENTUM": np.arange(n), "ENERGY": np.arange(n),
                  "STRESS": root},
        console_interval=100 * dt,
        info=dict(E=E, nu=nu, rho=rho, h=hh[0], wave_speed=c, particles=n))


def _hex_edges(n_cols, n_rows, side, origin):
    """Unique edges (k, 2, 2) of a pointy-top hexagon tiling of n_cols x n_rows
    cells with the given side length; odd rows shifted by half a cell."""
    w = np.sqrt(3.0) * side
    edges = {}
    for r in range(n_rows):
        for c in range(n_cols):
            cx = origin[0] + w * (c + 0.5 + 0.5 * (r % 2))
            cy = origin[1] + side + 1.5 * side * r
            ang = np.deg2rad(30.0 + 60.0 * np.arange(6))
            vx = cx + side * np.cos(ang)
            vy = cy + side * np.sin(ang)
            for k in range(6):
                a = (vx[k], vy[k])
                b = (vx[(k + 1) % 6], vy[(k + 1) % 6])
                key = tuple(sorted([(round(a[0] / side, 6), round(a[1] / side, 6)),
                                    (round(b[0] / side, 6), round(b[1] / side, 6))]))
                edges[key] = (a, b)
    return np.array(list(edges.values()))


def _segment_distance(px, py, edges):
    """(m,) points x (k,) segments -> (m, k) distances."""
    ax = edges[:, 0, 0][None, :]
    ay = edges[:, 0, 1][None, :]
    bx = edges[:, 1, 0][None, :]
    by = edges[:, 1, 1][None, :]
    dx, dy = bx - ax, by - ay
    t = ((px[:, None] - ax) * dx + (py[:, None] - ay) * dy) / (dx * dx + dy * dy)
    t = np.clip(t, 0.0, 1.0)
    qx = ax + t * dx - px[:, None]
    qy = ay + t * dy - py[:, None]
    return np.sqrt(qx * qx + qy * qy)


def honeycomb_crush(cells=(512, 512, 32), h=1.0e-4, hex_cells=(6, 6), z_layers=2, side_cells=None,
                    wall_cells=5.75, jitter=0.10, seed=1, E=70.0e9, nu=0.33, rho=2700.0, yield_stress=250.0e6,
                    hardening=1.0e9, platen_speed=10.0, platen_cells=2, steps=2000):
    """J2 honeycomb crush: a hexagonal honeycomb (hex_cells in x-y, extruded
    along z in z_layers layers) with walls wall_cells * h thick, each wall's
    thickness jittered uniformly by +-jitter from ``seed``; 2x2x2 particles per
    grid cell whose centre lies inside a wall.  Bottom support: node rows
    y < 1.5 h fixed in y (reaction measured).  Top platen: a slab of
    displacement-controlled particles in contact with the top walls, driven
    at -platen_speed in y from t = 0.

    Hardening maps the linear modulus H onto the reference's saturating
    law dR = C0 (sqrt(2/3) C1 - R) de_p
    (ConstitutiveRelation_calculateIncrement_VonMisesHardening.cpp:87-88):
    C1 = 1e3 sigma_y puts saturation far outside the working range and
    C0 = H / C1 makes the initial slope C0 * C1 = H.

    The default wall thickness, 5.75 h, is what puts 6 x 6 hexagonal cells of
    the widest side that fits a 512 x 512 x 32 grid at 8.0 M particles
    (4 h walls give 5.7 M)."""
    rng = np.random.default_rng(seed)
    cells = np.asarray(cells, dtype=np.int32)
    length = cells * h
    hh = length / cells
    nx, ny, nz = (int(c) for c in cells)
    if side_cells is None:
        # the widest honeycomb that fits x, with a 4-cell margin either side
        side_cells = (nx - 8) / (np.sqrt(3.0) * (hex_cells[0] + 0.5)) - 0.5
    side = side_cells * hh[0]
    origin = (4.0 * hh[0], 1.0 * hh[1] + 0.5 * wall_cells * hh[1] - 0.5 * side)
    edges = _hex_edges(hex_cells[0], hex_cells[1], side, origin)
    # keep the structure's lowest vertex at the bottom of cell row 1
    ymin = edges[:, :, 1].min()
    edges[:, :, 1] += (1.0 * hh[1] + 0.5 * wall_cells * hh[1]) - ymin
    # in-plane candidate cells near any wall
    ex0, ex1 = edges[:, :, 0].min() - wall_cells * hh[0], edges[:, :, 0].max() + wall_cells * hh[0]
    ey0, ey1 = edges[:, :, 1].min() - wall_cells * hh[1], edges[:, :, 1].max() + wall_cells * hh[1]
    cx0, cx1 = max(0, int(ex0 / hh[0])), min(nx - 1, int(ex1 / hh[0]) + 1)
    cy0, cy1 = max(1, int(ey0 / hh[1])), min(ny - 1, int(ey1 / hh[1]) + 1)
    gy, gx = np.meshgrid(np.arange(cy0, cy1 + 1), np.arange(cx0, cx1 + 1), indexing="ij")
    gx = gx.ravel()
    gy = gy.ravel()
    sub2 = np.array([[0.25, 0.25], [0.75, 0.25], [0.25, 0.75], [0.75, 0.75]])
    z0, z1 = 1, nz - 1  # body over z cells [1, nz-2]
    z_edges = np.linspace(z0, z1, z_layers + 1).round().astype(int)
    cell_rows = []
    for layer in range(z_layers):
        thick = wall_cells * hh[0] * (1.0 + rng.uniform(-jitter, jitter, size=len(edges)))
        inside_any = np.zeros((len(gx), 4), bool)
        for s in range(4):
            px = (gx + sub2[s, 0]) * hh[0]
            py = (gy + sub2[s, 1]) * hh[1]
            for k0 in range(0, len(px), 65536):
                sl = slice(k0, k0 + 65536)
                d = _segment_distance(px[sl], py[sl], edges)
                inside_any[sl, s] = (d <= 0.5 * thick[None, :]).any(axis=1)
        for cz in range(z_edges[layer], z_edges[layer + 1]):
            cell_rows.append((cz, inside_any))
    # emit particles cell by cell: z, then y, then x over cells; sub-order z, y, x
    xs = []
    for cz, inside in cell_rows:
        keep = inside.any(axis=1)
        kx, ky, km = gx[keep], gy[keep], inside[keep]
        # sub positions in (oz, oy, ox) order with ox fastest
        pts = []
        for oz in (0.25, 0.75):
            for s_y in (0, 1):
                for s_x in (0, 1):
                    s = s_y * 2 + s_x
                    pts.append((s, oz))
        px = np.stack([(kx + sub2[s, 0]) * hh[0] for s, _ in pts], axis=1)
        py = np.stack([(ky + sub2[s, 1]) * hh[1] for s, _ in pts], axis=1)
        pz = np.stack([np.full(len(kx), (cz + oz) * hh[2]) for _, oz in pts], axis=1)
        pm = np.stack([km[:, s] for s, _ in pts], axis=1)
        xs.append(np.stack([px[pm], py[pm], pz[pm]]))
    x_body = np.concatenate(xs, axis=1)
    n_body = x_body.shape[1]
    # platen: particles over the honeycomb's x extent, in the cells right above its top
    top = x_body[1].max()
    py0 = int(np.floor(top / hh[1])) + 1
    pcx0 = int(np.floor(x_body[0].min() / hh[0]))
    pcx1 = int(np.floor(x_body[0].max() / hh[0]))
    pc = np.array([[ix, iy, iz] for iz in range(z0, z1) for iy in range(py0, py0 + platen_cells)
                   for ix in range(pcx0, pcx1 + 1)], dtype=np.int64)
    x_platen = lattice_in_cells(pc, hh)
    if py0 + platen_cells >= ny:
        raise ValueError("grid too short for the platen")
    x = np.concatenate([x_body, x_platen], axis=1)
    n = x.shape[1]
    v = np.zeros((3, n))
    offset = hh[0] / 2.0
    fields = default_fields(x, v, offset ** 3, rho)
    c1 = 1.0e3 * yield_stress
    c0 = hardening / c1
    mats = [dict(type=VONMISESHARDENING, density=rho, elastic_modulus=E, poisson_ratio=nu,
                 yield_stress=yield_stress, hardening_c0=c0, hardening_c1=c1)]
    nodes = node_positions(length, cells)
    fixed = np.where(np.abs(nodes[1] - 0.0) < 1.5 * hh[1], AXIS_Y, 0).astype(np.uint8)
    platen = np.arange(n_body, n, dtype=np.int64)
    c = np.sqrt(E / rho)
    dt = 0.1 * hh[0] / c
    return Scene(
        name=f"honeycomb_crush_{nx}x{ny}x{nz}", length=length, cells=cells, fixed=fixed,
        friction=np.zeros(fixed.shape), materials=mats, fields=fields, material=np.zeros(n, np.int32),
        dt=dt, time_end=steps * dt,
        dc_index=platen, dc_axes=np.full(len(platen), AXIS_Y, np.uint8),
        dc_mult=np.full(len(platen), -1.0, np.float32),
        monitors={"DISPLACEMENT": platen, "MOMENTUM": np.arange(n_body, dtype=np.int64),
                  "ENERGY": np.arange(n_body, dtype=np.int64)},
        timeline_t=np.array([0.0, 10.0]), timeline_v=np.array([[0.0, platen_speed, 0.0], [0.0, platen_speed, 0.0]]),
        console_interval=100 * dt,
        info=dict(E=E, nu=nu, rho=rho, yield_stress=yield_stress, hardening=hardening, c0=c0, c1=c1,
                  h=hh[0], side_cells=side_cells, wall_cells=wall_cells, seed=seed, particles=n,
                  body_particles=n_body, platen_particles=len(platen), hex_cells=hex_cells, z_layers=z_layers,
                  wave_speed=c))


def half_ring_impact(h=5.0e-5, depth_cells=91, foam_cells=80, porosity=0.5, impact_speed=20.0, seed=4,
                     platen_cells=2, steps=1000):
    """C4: the half ring of
    PhysicsEngine_initializeWorld_Classic_HalfRing_Xiang_PlainStress.cpp:4-195
    crushed by a foam impactor.  The ring (inner diameter 47.9 mm, wall
    1.48 mm, steel, von Mises hardening with the reference's constants) is cut
    at its symmetry plane: it stands on the y support (nodes y < 1.5 h fixed in
    y, reaction measured), its cut face lies on the nodes x < 1.5 h, fixed in
    x, and every node is fixed in z (the reference's plane-stress set-up,
    here extruded over depth_cells cells).  Shifts: +h in x as in the
    reference; +h in y (the reference's +2h leaves the ring's lowest cells
    one node row above the fixed rows y < 1.5 h, so it would meet the support
    only after moving a cell).  On top of the ring, in contact with it, a block of aluminium
    foam (the reference's second material) foam_cells tall over 0.95 of the
    ring's width, with spherical voids (radii lognormal, mu = ln 3h,
    sigma = 0.3, from ``seed``) up to ``porosity``, moves down at
    impact_speed; its top platen_cells layers are displacement-controlled at
    that speed.  2x2x2 particles per solid cell; the default h gives about
    48 M particles."""
    rng = np.random.default_rng(seed)
    r_in = 0.5 * 0.0479
    r_out = r_in + 0.00148
    ring_cx, ring_cy = 1.0 * h, 1.0 * h + r_out
    nz = int(depth_cells) + 2
    ring_top = ring_cy + r_out
    nx = int(np.ceil((ring_cx + r_out) / h)) + 4
    ny_ring = int(np.ceil(ring_top / h)) + 1
    # ring: cells whose centre lies in the half annulus (x >= the cut), same in every z layer
    gy, gx = np.meshgrid(np.arange(ny_ring), np.arange(nx), indexing="ij")
    cxp, cyp = (gx + 0.5) * h, (gy + 0.5) * h
    rr = np.hypot(cxp - ring_cx, cyp - ring_cy)
    ring2d = (rr >= r_in) & (rr <= r_out) & (cxp >= ring_cx)
    # the foam's first cell row sits on the ring's top row: the two share that
    # row of nodes, so they are in contact from the start
    fy0 = int(np.nonzero(ring2d)[0].max()) + 1
    fy1 = fy0 + int(foam_cells)                   # foam rows [fy0, fy1), platen on top
    ny = fy1 + platen_cells + 3
    cells = np.array([nx, ny, nz], dtype=np.int32)
    hh = np.full(3, h)
    length = cells * hh
    # foam: a block of cells, voids carved by bounding box
    fx0 = int(np.floor(ring_cx / h))
    fx1 = fx0 + int(np.floor(0.95 * r_out / h))
    body = np.array([fx1 - fx0, fy1 - fy0, nz - 2])
    solid = np.ones((body[2], body[1], body[0]), bool)          # [z, y, x]
    target = porosity * solid.size
    removed = 0
    lo = np.array([fx0, fy0, 1], np.float64) * h
    while removed < target:
        r = float(np.exp(rng.normal(np.log(3.0 * h), 0.3)))
        c = lo + rng.uniform(0, 1, 3) * body * h
        a0 = np.clip(np.floor((c - r - lo) / h).astype(int), 0, body - 1)
        a1 = np.clip(np.floor((c + r - lo) / h).astype(int) + 1, 0, body)
        sz, sy, sx = (slice(a0[2], a1[2]), slice(a0[1], a1[1]), slice(a0[0], a1[0]))
        zc = (np.arange(a0[2], a1[2]) + 0.5) * h + lo[2] - c[2]
        yc = (np.arange(a0[1], a1[1]) + 0.5) * h + lo[1] - c[1]
        xc = (np.arange(a0[0], a1[0]) + 0.5) * h + lo[0] - c[0]
        ball = (zc[:, None, None] ** 2 + yc[None, :, None] ** 2 + xc[None, None, :] ** 2) <= r * r
        sub = solid[sz, sy, sx]
        removed += int((sub & ball).sum())
        sub &= ~ball
    # cells of each body in emission order (z, then y, then x)
    ry, rx = np.nonzero(ring2d)
    ring_cells = np.concatenate([np.stack([rx, ry, np.full(len(rx), z)], 1) for z in range(1, nz - 1)])
    fz, fyy, fxx = np.nonzero(solid)
    foam_cells_idx = np.stack([fxx + fx0, fyy + fy0, fz + 1], 1)
    pz, py_, px_ = np.meshgrid(np.arange(1, nz - 1), np.arange(fy1, fy1 + platen_cells), np.arange(fx0, fx1),
                               indexing="ij")
    platen_idx = np.stack([px_.ravel(), py_.ravel(), pz.ravel()], 1)
    x = np.concatenate([lattice_in_cells(ring_cells, hh), lattice_in_cells(foam_cells_idx, hh),
                        lattice_in_cells(platen_idx, hh)], 1)
    n_ring, n_foam, n_platen = 8 * len(ring_cells), 8 * len(foam_cells_idx), 8 * len(platen_idx)
    n = x.shape[1]
    v = np.zeros((3, n))
    v[1, n_ring:] = -impact_speed
    steel = dict(type=VONMISESHARDENING, density=7800.0, elastic_modulus=210.0e9, poisson_ratio=0.3,
                 yield_stress=310.0e6, hardening_c0=4.0, hardening_c1=150.0e6)
    aluminium = dict(type=VONMISESHARDENING, density=2760.0, elastic_modulus=70.0e9, poisson_ratio=0.3,
                     yield_stress=190.0e6, hardening_c0=30.0, hardening_c1=50.0e6)
    material = np.zeros(n, np.int32)
    material[n_ring:] = 1
    fields = default_fields(x, v, (h / 2.0) ** 3, 1.0)
    fields["MASS"] = np.where(material == 0, steel["density"], aluminium["density"]) * fields["VOLUME_INITIAL"]
    nodes = node_positions(length, cells)
    fixed = (np.where(nodes[0] < 1.5 * h, AXIS_X, 0) | np.where(nodes[1] < 1.5 * h, AXIS_Y, 0) | AXIS_Z)
    fixed = fixed.astype(np.uint8)
    platen = np.arange(n_ring + n_foam, n, dtype=np.int64)
    c = np.sqrt(steel["elastic_modulus"] / steel["density"])
    dt = 0.1 * h / c
    return Scene(
        name=f"half_ring_impact_{nx}x{ny}x{nz}", length=length, cells=cells, fixed=fixed,
        friction=np.zeros(fixed.shape), materials=[steel, aluminium], fields=fields, material=material,
        dt=dt, time_end=steps * dt,
        dc_index=platen, dc_axes=np.full(len(platen), AXIS_X | AXIS_Y | AXIS_Z, np.uint8),
        dc_mult=np.full(len(platen), -1.0, np.float32),
        monitors={"DISPLACEMENT": platen, "MOMENTUM": np.arange(n_ring + n_foam, dtype=np.int64),
                  "ENERGY": np.arange(n_ring, dtype=np.int64)},
        timeline_t=np.array([0.0, 10.0]),
        timeline_v=np.array([[0.0, impact_speed, 0.0], [0.0, impact_speed, 0.0]]),
        console_interval=100 * dt,
        info=dict(h=h, particles=n, ring_particles=n_ring, foam_particles=n_foam, platen_particles=n_platen,
                  porosity=removed / solid.size, seed=seed, impact_speed=impact_speed, wave_speed=c))


def random_block(seed=0, cells=(6, 5, 4), n=300, h=1.0e-3, features=True):
    """Small randomized scene exercising every branch of the step: three
    materials (elastic, perfectly plastic, hardening) plus an unknown type and
    a missing material, random stresses near yield, displacement-controlled
    particles on all axes with mixed multipliers, fixed nodes with friction,
    damping, external forces, a two-point timeline and particles that touch
    the grid edge or lie off it."""
    rng = np.random.default_rng(seed)
    cells = np.asarray(cells, dtype=np.int32)
    hh = np.full(3, h)
    length = cells * hh
    x = rng.uniform(-0.6, 1.0, size=(3, n)) * length[:, None] + 0.3 * length[:, None] * rng.uniform(0, 1, (3, n))
    x = np.clip(x, -0.4 * h, length[:, None] + 0.4 * h)
    if features:
        # a few exactly on node planes and a few off the grid
        x[0, :5] = np.round(x[0, :5] / h) * h
        x[:, 5] = length + 3.0 * h
        x[:, 6] = -3.0 * h
    v = rng.normal(0.0, 2.0, size=(3, n))
    E = 5.0e9
    rho = 1200.0
    vol = (0.5 * h) ** 3
    fields = default_fields(x, v, vol, rho)
    fields["MASS"] = fields["MASS"] * rng.uniform(0.5, 1.5, n)
    sig = rng.normal(0.0, 30.0e6, size=(6, n))
    fields["STRESS"] = sig
    fields["BACKSTRESS"] = np.abs(rng.normal(0.0, 5.0e6, n))
    fields["STRAIN"] = rng.normal(0.0, 1e-3, size=(6, n))
    F = np.zeros((9, n))
    F[0] = F[4] = F[8] = 1.0
    F += rng.normal(0.0, 1e-3, size=(9, n))
    fields["DEFORMATION_GRADIENT"] = F
    if features:
        fields["FORCE_EXTERNAL"] = np.where(rng.uniform(size=(3, n)) < 0.3, rng.normal(0.0, 1e-3, (3, n)), 0.0)
    mats = [
        dict(type=ELASTIC, density=rho, elastic_modulus=E, poisson_ratio=0.3),
        dict(type=PLASTIC, density=rho, elastic_modulus=E, poisson_ratio=0.25, yield_stress=40.0e6),
        dict(type=VONMISESHARDENING, density=rho, elastic_modulus=E, poisson_ratio=0.33, yield_stress=35.0e6,
             hardening_c0=15.0, hardening_c1=60.0e6),
        dict(type=7, density=rho, elastic_modulus=E, poisson_ratio=0.2, yield_stress=45.0e6),
    ]
    material = rng.integers(0, 4, n).astype(np.int32)
    if features:
        material[rng.uniform(size=n) < 0.03] = -1
    nodes = node_positions(length, cells)
    fixed = np.zeros(nodes.shape[1], np.uint8)
    friction = np.zeros(nodes.shape[1])
    if features:
        fixed[nodes[0] < 1.5 * h] |= AXIS_X
        fixed[nodes[1] < 0.5 * h] |= AXIS_Y
        fixed[nodes[2] > length[2] - 0.5 * h] |= AXIS_Z
        friction[:] = rng.uniform(0.0, 0.6, nodes.shape[1])
    dc = np.sort(rng.choice(n, size=n // 10, replace=False)).astype(np.int64) if features else np.zeros(0, np.int64)
    rng.shuffle(dc)  # list order differs from index order
    dc_axes = rng.integers(1, 8, len(dc)).astype(np.uint8)
    dc_mult = rng.choice([-1.0, 1.0, 0.5], len(dc)).astype(np.float32)
    c = np.sqrt(E / rho)
    dt = 0.1 * h / c
    return Scene(
        name=f"random_block_{seed}", length=length, cells=cells, fixed=fixed, friction=friction,
        materials=mats, fields=fields, material=material, dt=dt, time_end=50 * dt,
        dc_index=dc, dc_axes=dc_axes, dc_mult=dc_mult,
        monitors={"DISPLACEMENT": dc[: max(1, len(dc) // 2)] if len(dc) else np.zeros(0, np.int64),
                  "MOMENTUM": np.arange(n), "ENERGY": np.arange(0, n, 2), "STRESS": np.arange(0, n, 3)},
        timeline_t=np.array([0.0, 5 * dt, 1.0]) if features else np.zeros(0),
        timeline_v=np.array([[0.0, 0.0, 0.0], [0.3, -2.0, 0.1], [0.3, -2.0, 0.1]]) if features else np.zeros((0, 3)),
        damping=0.05 if features else 0.0,
        console_interval=10 * dt)


def crop_x(scene, cells_x):
    """The part of a scene with x < cells_x * h on a grid cut to cells_x cells
    in x (node rows beyond it dropped); lists are filtered and re-indexed."""
    h = scene.h
    keep = np.nonzero(scene.fields["POSITION"][0] < cells_x * h[0])[0]
    remap = -np.ones(scene.n, np.int64)
    remap[keep] = np.arange(len(keep))
    cells = np.array([cells_x, scene.cells[1], scene.cells[2]], np.int32)
    length = np.array([cells_x * h[0], scene.length[1], scene.length[2]])
    nn_old = np.asarray(scene.cells) + 1
    nn = cells + 1
    iz, iy, ix = np.meshgrid(np.arange(nn[2]), np.arange(nn[1]), np.arange(nn[0]), indexing="ij")
    old = (ix + nn_old[0] * (iy + nn_old[1] * iz)).ravel()

    def sub(idx):
        r = remap[np.asarray(idx, np.int64)]
        return r[r >= 0]

    dc_keep = remap[scene.dc_index] >= 0
    return Scene(
        name=scene.name + f"_x{cells_x}", length=length, cells=cells, fixed=scene.fixed[old],
        friction=scene.friction[old], materials=scene.materials,
        fields={k: np.ascontiguousarray(v[..., keep]) for k, v in scene.fields.items()},
        material=scene.material[keep], dt=scene.dt, time_end=scene.time_end,
        dc_index=remap[scene.dc_index][dc_keep], dc_axes=scene.dc_axes[dc_keep], dc_mult=scene.dc_mult[dc_keep],
        monitors={k: sub(v) for k, v in scene.monitors.items()}, timeline_t=scene.timeline_t,
        timeline_v=scene.timeline_v, damping=scene.damping, mass_min=scene.mass_min, time0=scene.time0,
        console_interval=scene.console_interval, info=dict(scene.info, particles=len(keep)))


def graded_foam(seed, h=1.0e-4, n_target=None, E=10.0e9, nu=0.3, rho=1000.0, yield_stress=100.0e6,
                platen_speed=5.0, steps=500):
    """One graded-foam instance (the batched workload): a cuboid of 2x2x2
    particles per cell with spherical voids whose radii are lognormal
    (mu = ln 3h, sigma = 0.3) and whose number density grows linearly along
    z (three times denser at the back face than at the front), across the
    loading axis y; porosity ~ U[0.3, 0.7] and particle count ~ log-uniform
    [1e4, 2e5], both drawn from ``seed``.  J2 perfectly plastic.  Fixed-y
    support below, displacement-controlled platen on top at platen_speed."""
    rng = np.random.default_rng(seed)
    porosity = rng.uniform(0.3, 0.7)
    if n_target is None:
        n_target = int(np.exp(rng.uniform(np.log(1.0e4), np.log(2.0e5))))
    solid_cells = n_target / 8.0 / (1.0 - porosity)
    s = max(6, int(round((solid_cells / 1.5) ** (1.0 / 3.0))))
    body = np.array([s, int(round(1.5 * s)), s])  # y is the loading axis
    margin = 2
    cells = body + np.array([2 * margin, 2 * margin + 3, 2 * margin])
    hh = np.full(3, h)
    length = cells * hh
    iz, iy, ix = np.meshgrid(np.arange(body[2]), np.arange(body[1]), np.arange(body[0]), indexing="ij")
    cidx = np.stack([ix.ravel(), iy.ravel(), iz.ravel()], 1) + np.array([margin, 1, margin])
    x = lattice_in_cells(cidx, hh)
    lo = np.array([margin, 1, margin]) * h
    hi = lo + body * h
    keep = np.ones(x.shape[1], bool)
    # voids: add until the porosity is reached; centre density linear in z (1 at the front face z = lo,
    # 3 at the back), across the loading axis y.  Particles come 8 per cell, cells x fastest: a void
    # only tests the cells of its bounding box.
    target_removed = porosity * x.shape[1]
    removed = 0
    sub8 = np.arange(8)
    for _ in range(100000):
        if removed >= target_removed:
            break
        r = np.exp(rng.normal(np.log(3.0 * h), 0.3))
        while True:
            c = lo + rng.uniform(0, 1, 3) * (hi - lo)
            if rng.uniform() < (1.0 + 2.0 * (c[2] - lo[2]) / (hi[2] - lo[2])) / 3.0:
                break
        a0 = np.clip(np.floor((c - r - lo) / h).astype(int) - 1, 0, body - 1)
        a1 = np.clip(np.floor((c + r - lo) / h).astype(int) + 1, 0, body - 1)
        bz, by, bx = np.meshgrid(np.arange(a0[2], a1[2] + 1), np.arange(a0[1], a1[1] + 1),
                                 np.arange(a0[0], a1[0] + 1), indexing="ij")
        cj = (bx + body[0] * (by + body[1] * bz)).ravel()
        cand = (8 * cj[:, None] + sub8[None, :]).ravel()
        d2 = np.sum((x[:, cand] - c[:, None]) ** 2, 0)
        hit = cand[(d2 <= r * r) & keep[cand]]
        keep[hit] = False
        removed += len(hit)
    x = np.ascontiguousarray(x[:, keep])
    n_body = x.shape[1]
    # platen: one cell layer right above the body
    py = 1 + body[1]
    pc = np.array([[i, py, k] for k in range(margin, margin + body[2]) for i in range(margin, margin + body[0])])
    xp = lattice_in_cells(pc, hh)
    x = np.concatenate([x, xp], 1)
    n = x.shape[1]
    fields = default_fields(x, np.zeros((3, n)), (h / 2) ** 3, rho)
    mats = [dict(type=PLASTIC, density=rho, elastic_modulus=E, poisson_ratio=nu, yield_stress=yield_stress)]
    nodes = node_positions(length, cells)
    fixed = np.where(nodes[1] < 1.5 * h, AXIS_Y, 0).astype(np.uint8)
    platen = np.arange(n_body, n, dtype=np.int64)
    c = np.sqrt(E / rho)
    dt = 0.1 * h / c
    return Scene(
        name=f"graded_foam_{seed}", length=length, cells=cells.astype(np.int32), fixed=fixed,
        friction=np.zeros(fixed.shape), materials=mats, fields=fields, material=np.zeros(n, np.int32), dt=dt,
        time_end=steps * dt, dc_index=platen, dc_axes=np.full(len(platen), AXIS_Y, np.uint8),
        dc_mult=np.full(len(platen), -1.0, np.float32),
        monitors={"DISPLACEMENT": platen, "MOMENTUM": np.arange(n_body, dtype=np.int64),
                  "ENERGY": np.arange(n_body, dtype=np.int64)},
        timeline_t=np.array([0.0, 10.0]), timeline_v=np.array([[0.0, platen_speed, 0.0], [0.0, platen_speed, 0.0]]),
        console_interval=50 * dt,
        info=dict(porosity=porosity, particles=n, body_particles=n_body, seed=seed, E=E, nu=nu, rho=rho,
                  yield_stress=yield_stress))
