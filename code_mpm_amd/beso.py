"""Bi-directional evolutionary structural optimisation (BESO) over the MPM
timestep: the reference's outer workload (Main.cpp:127-336), run as a batched
client of the engine.

  - `Canvas2D`: the voxel canvas of Canvas2D_CC (Canvas2D_CC.h:9-81,
    Canvas2D_CC.cpp): voxel flags, objective, ESO opacity, the draw / cut /
    flag primitives, and the objective smoothing filter;
  - `eso_scene`: the world PhysicsEngine::initializeWorld_Classic_ESO builds
    from a canvas (PhysicsEngine_initializeWorld_Classic_ESO.cpp:4-307): one
    particle per active voxel, volume scaled by the voxel's opacity, a
    displacement-controlled load patch, a fixed top edge;
  - `BESO`: the iteration loop of Main.cpp:204-333 for several designs at
    once.  Each iteration builds every design's world, runs them to their end
    time side by side as one batch of instances (one set of kernel launches
    per step for all of them), folds the particles' maximum-history strain
    energy (the objective of Main.cpp:24-48, computed on the device by the
    step) into the voxels, smooths it (`mpm_eso_filter`, on the device),
    then soft-kills and restores voxels by the reference's ratios.

The simulation and the filter are injectable callables, so the tests run
the same driver over a CPU checker; the product runs both on the GPU and
fails loudly without the HIP library.  Sorting: the reference calls std::sort, whose order of
equal objectives is unspecified; here equal objectives keep ascending voxel
index (a stable sort), and NaN objectives sort last.
"""
import math

import numpy as np

from . import scenes

ELASTIC = 0


class Canvas2D:
    """Canvas2D_CC: an nx x ny voxel canvas, index x + y * nx
    (Canvas2D_CC.h:61-76), voxel centre (i + 0.5) * offset."""

    def __init__(self, size, offset):
        self.d_Offset = float(offset)
        self.d2_Size = (float(size[0]), float(size[1]))
        # Canvas2D_CC.cpp:9-11: floor(size / offset)
        self.nx = int(math.floor(self.d2_Size[0] / self.d_Offset))
        self.ny = int(math.floor(self.d2_Size[1] / self.d_Offset))
        nv = self.nx * self.ny
        ids = np.arange(nv, dtype=np.int64)
        ix, iy = ids % self.nx, ids // self.nx
        # Canvas2D_CC.cpp:22: dvec2(index) * offset + 0.5 * dvec2(offset)
        self.pos = np.stack([ix.astype(np.float64) * self.d_Offset + 0.5 * self.d_Offset,
                             iy.astype(np.float64) * self.d_Offset + 0.5 * self.d_Offset])
        self.active = np.zeros(nv, bool)
        self.surface = np.zeros(nv, bool)
        self.load = np.zeros(nv, bool)
        self.support = np.zeros(nv, bool)
        self.eso = np.zeros(nv, bool)
        self.redundant = np.zeros(nv, bool)
        self.objective = np.zeros(nv)
        self.opacity = np.ones(nv)   # Voxel_ST::d_ESO_Opacity default 1.0

    @property
    def n_voxels(self):
        return self.nx * self.ny

    # -- geometry ---------------------------------------------------------------
    def _local(self, center, rotation):
        """Voxel centres in the frame of a shape centred at `center` and
        rotated by `rotation` (the translate-then-rotate of Canvas2D_CC.cpp:149-159).
        With rotation 0 this is the exact difference p - c, as glm's matrix
        product gives; other rotations are restated from the mathematics."""
        dx = self.pos[0] - float(center[0])
        dy = self.pos[1] - float(center[1])
        if rotation == 0.0:
            return dx, dy
        c, s = math.cos(-rotation), math.sin(-rotation)
        return c * dx - s * dy, s * dx + c * dy

    def _in_rect(self, center, size, rotation):
        lx, ly = self._local(center, rotation)
        hx, hy = 0.5 * float(size[0]), 0.5 * float(size[1])
        return ~((lx < -hx) | (hx < lx)) & ~(np.abs(ly) > hy)

    def _dist(self, center):
        dx = self.pos[0] - float(center[0])
        dy = self.pos[1] - float(center[1])
        return np.sqrt(dx * dx + dy * dy)   # glm::length

    def drawCircle(self, center, radius):
        """Canvas2D_CC.cpp:42-56."""
        self.active |= ~(self._dist(center) > radius)

    def cutCircle(self, center, radius):
        """Canvas2D_CC.cpp:195-208."""
        self.active &= self._dist(center) > radius

    def drawRing(self, center, r_outer, r_inner):
        """Canvas2D_CC.cpp:58-84: voxels between the radii (half a voxel in
        from each), with the surface flag of the ring's edges kept only on
        voxels that were not already interior."""
        d = self._dist(center)
        inside = ~((d < r_inner + 0.5 * self.d_Offset) | (r_outer - 0.5 * self.d_Offset < d))
        surf = (np.abs(d - r_outer) < self.d_Offset) | (np.abs(d - r_inner) < self.d_Offset)
        surf = np.where(surf & self.active, self.surface, surf)
        self.surface = np.where(inside, surf, self.surface)
        self.active |= inside

    def drawRectangle(self, center, size, rotation=0.0):
        """Canvas2D_CC.cpp:86-112."""
        self.active |= self._in_rect(center, size, rotation)

    def drawPrism(self, center, size, extra, rotation=0.0):
        """Canvas2D_CC.cpp:114-140: a rectangle whose height shrinks
        linearly along x by `extra` over the half width."""
        lx, ly = self._local(center, rotation)
        hx, hy = 0.5 * float(size[0]), 0.5 * float(size[1])
        t = 0.5 * extra * (2.0 * lx / float(size[0]))
        ok = ~((lx < -hx) | (hx < lx)) & ~((ly < -hy + t) | (hy - t < ly))
        self.active |= ok

    def cutRectangle(self, center, size, rotation=0.0):
        """Canvas2D_CC.cpp:210-236."""
        self.active &= ~self._in_rect(center, size, rotation)

    def setESORectangle(self, center, size, rotation, flag):
        """Canvas2D_CC.cpp:238-262."""
        self.eso = np.where(self._in_rect(center, size, rotation), bool(flag), self.eso)

    def setLoadRectangle(self, center, size, rotation, flag):
        self.load = np.where(self._in_rect(center, size, rotation), bool(flag), self.load)

    def setSupportRectangle(self, center, size, rotation, flag):
        self.support = np.where(self._in_rect(center, size, rotation), bool(flag), self.support)

    # -- queries -----------------------------------------------------------------
    def getVoxels_Active(self, state=True):
        """Indices of voxels whose b_Active equals state, ascending (Canvas2D_CC.cpp:264-277)."""
        return np.flatnonzero(self.active == bool(state))

    def getVoxels_ESO(self, state=True):
        return np.flatnonzero(self.eso == bool(state))

    def getCount_Active(self, state=True):
        return int((self.active == bool(state)).sum())

    def getCount_ESO(self, state=True):
        return int((self.eso == bool(state)).sum())

    def getObjective_Sum(self):
        """Sum of the active voxels' objectives, added in index order (Canvas2D_CC.cpp:343-357)."""
        a = self.objective[self.active]
        return float(np.cumsum(a)[-1]) if a.size else 0.0

    def filterObjective_Smooth(self, radius, device=0, smooth=None):
        """Canvas2D_CC.cpp:302-341 on the GPU (mpm_eso_filter); `smooth`
        replaces it with a function of (active, eso, objective, nx, ny,
        offset, radius)."""
        if smooth is None:
            from . import capi
            self.objective = capi.eso_filter(self.active, self.eso, self.objective, self.nx, self.ny,
                                             self.d_Offset, radius, device)
        else:
            self.objective = smooth(self.active, self.eso, self.objective, self.nx, self.ny, self.d_Offset, radius)

    def copy(self):
        c = Canvas2D.__new__(Canvas2D)
        c.__dict__.update({k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in self.__dict__.items()})
        return c


def l_problem(offset=0.0001, size=(0.050, 0.050)):
    """The L-shaped domain of Main.cpp:127-145: a 30 mm square with an
    18 mm square cut from its top right corner, and a 6-voxel load patch at
    the tip of the lower arm (its surroundings, 24 voxels wide, kept out of
    the optimisation).  All dimensions scale with the canvas size."""
    sx, sy = float(size[0]), float(size[1])
    k = sx / 0.050
    a, b = 0.030 * k, 0.018 * k
    cv = Canvas2D((sx, sy), offset)
    cv.drawRectangle((0.5 * sx, sy - 0.5 * a), (a, a), 0.0)
    cv.setESORectangle((0.5 * sx, sy - 0.5 * a), (a, a), 0.0, True)
    cv.cutRectangle((0.5 * sx + 0.5 * a - 0.5 * b, sy - 0.5 * b), (b, b), 0.0)
    cv.setESORectangle((0.5 * sx + 0.5 * a - 0.5 * b, sy - 0.5 * b), (b, b), 0.0, False)
    F = (0.5 * sx + 0.5 * a + 3 * offset, sy - a + 0.006 * k)
    cv.drawRectangle(F, (6.1 * offset, 6.1 * offset), 0.0)
    cv.setESORectangle(F, (24.1 * offset, 24.1 * offset), 0.0, False)
    cv.setLoadRectangle(F, (6.1 * offset, 6.1 * offset), 0.0, True)
    return cv


def eso_scene(canvas, dt_max=1.0 / 4.0 * 5.0e-8, time_loading=2.0 * 1.0e-4, displacement_max=0.001 * 1.0e-3,
              damping=0.1):
    """PhysicsEngine::initializeWorld_Classic_ESO
    (PhysicsEngine_initializeWorld_Classic_ESO.cpp:4-307) for a canvas: a
    grid of 2-voxel cells one cell thick, nodes on the top edge fixed in x
    and y, every node fixed in z; one elastic particle per active voxel at
    the voxel centre shifted by one cell in x, with volume offset^3 times the
    voxel's opacity; load voxels controlled in y with multiplier -1 and
    support voxels in y with multiplier 0; the load ramps to a peak speed at
    half the loading time and back to rest.  The defaults are the
    reference's; tests shorten the loading time."""
    off = canvas.d_Offset
    Lg = np.array([canvas.d2_Size[0], canvas.d2_Size[1], 1.0 * 2.0 * off / 1.0])
    cells = np.floor(Lg / (2.0 * off)).astype(np.int32)
    h = Lg / cells
    nodes = scenes.node_positions(Lg, cells)
    tol = 0.1 * h[0]
    fixed = np.zeros(nodes.shape[1], np.uint8)
    top = np.abs(nodes[1] - Lg[1]) < tol
    fixed[top] |= 1 | 2
    fixed[np.abs(nodes[2] - 0.0) < 2.0 * Lg[2]] |= 4
    E = 10.0 * 1.0e9
    mat = dict(type=ELASTIC, density=10.0 * 40.0, elastic_modulus=E, poisson_ratio=0.3, yield_stress=0.001 * E,
               hardening_c0=35.0, hardening_c1=0.5 * (0.001 * E))
    ids = canvas.getVoxels_Active(True)
    n = ids.size
    x = np.stack([canvas.pos[0, ids] + h[0], canvas.pos[1, ids] + 0.0, np.full(n, 0.0 + 0.5 * Lg[2])])
    V0 = off * off * off * canvas.opacity[ids]
    fields = {
        "POSITION": x,
        "VELOCITY": np.zeros((3, n)),
        "MASS": mat["density"] * V0,
        "VOLUME": V0.copy(),
        "VOLUME_INITIAL": V0,
    }
    load = canvas.load[ids]
    support = canvas.support[ids] & ~load
    dc = np.flatnonzero(load | support)
    t_resp = 1.0 * time_loading
    speed = 2.0 * displacement_max / time_loading
    t_end = 1.0 * t_resp
    sc = scenes.Scene(
        name=f"eso_{canvas.nx}x{canvas.ny}", length=Lg, cells=cells, fixed=fixed, friction=np.zeros(nodes.shape[1]),
        materials=[mat], fields=fields, material=np.zeros(n, np.int32), dt=dt_max, time_end=t_end,
        dc_index=dc.astype(np.int64), dc_axes=np.full(dc.size, 2, np.uint8),
        dc_mult=np.where(load[dc], -1.0, 0.0).astype(np.float32),
        monitors={"ENERGY": np.arange(n, dtype=np.int64), "DISPLACEMENT": np.flatnonzero(load).astype(np.int64),
                  "MOMENTUM": np.arange(n, dtype=np.int64)},
        timeline_t=np.array([0.0, 0.5 * time_loading, 1.0 * time_loading]),
        timeline_v=np.array([[0.0, 0.0, 0.0], [0.0, speed, 0.0], [0.0, 0.0, 0.0]]),
        damping=damping, mass_min=0.0, console_interval=0.05 * t_end,
        info=dict(voxel_ids=ids))
    return sc


def script_e(x, precision):
    """Script(double, precision) of Definitions.h:46-60: scientific, explicit sign."""
    return f"{x:+.{precision}e}"


class Design:
    """One design's BESO state (Main.cpp:173-183)."""

    def __init__(self, canvas, ratio_redundant=0.05, ratio_rejection=0.05, ratio_increment=0.05,
                 opacity_min=1.0e-3, opacity_max=1.0, radius=0.0012):
        self.canvas = canvas
        self.ratio_redundant = ratio_redundant
        self.ratio_rejection = ratio_rejection
        self.ratio_increment = ratio_increment
        self.opacity_min = opacity_min
        self.opacity_max = opacity_max
        self.radius = radius
        self.iteration_local = 0
        self.removed = 0
        self.restored = 0
        self.log = []

    def fold_objective(self, voxel_ids, wmax):
        """Main.cpp:236-257: each particle's objective (its maximum-history
        strain energy over one particle, getObjective) halves into its voxel's."""
        obj = self.canvas.objective
        obj[voxel_ids] = 0.5 * obj[voxel_ids] + 0.5 * (np.asarray(wmax, np.float64) / 1.0)

    def update(self):
        """Main.cpp:262-314: sort the ESO voxels by objective, soft-kill the
        lowest with opacity above 0.8, restore the highest with opacity below
        0.8, each up to the redundancy budget; raise the rejection ratio when
        the design has settled or after 100 local iterations."""
        cv = self.canvas
        active = cv.getVoxels_Active(True)
        eso = cv.getVoxels_ESO(True)
        order = eso[np.argsort(cv.objective[eso], kind="stable")]
        n_eso = order.size
        budget = self.ratio_redundant * (1.0 - self.ratio_rejection) * n_eso
        removed = 0
        i = 0
        while i < self.ratio_rejection * n_eso:
            v = order[i]
            if cv.eso[v]:
                if cv.opacity[v] > 0.8:
                    cv.opacity[v] = self.opacity_min
                    removed += 1
                if removed > budget:
                    break
            i += 1
        restored = 0
        i = n_eso - 1
        while i >= 0 and i >= self.ratio_rejection * n_eso:
            v = order[i]
            if cv.eso[v]:
                if cv.opacity[v] < 0.8:
                    cv.opacity[v] = self.opacity_max
                    restored += 1
                if restored > budget:
                    break
            i -= 1
        if removed + restored < 0.002 * active.size or self.iteration_local == 100:
            self.ratio_rejection += self.ratio_increment
            self.iteration_local = 0
        else:
            self.iteration_local += 1
        self.removed, self.restored = removed, restored

    def log_line(self, iteration, total_energy):
        """The BESO log record of Main.cpp:316-333."""
        return (f"iIteration_Global: \t{iteration}\tiIteration_Local: \t{self.iteration_local}\t"
                f"dRatio_Rejection: \t{script_e(self.ratio_rejection, 3)}\tRemoved: \t{self.removed}\t"
                f"Restored: \t{self.restored}\tObjective_Sum: \t{script_e(self.canvas.getObjective_Sum(), 5)}\t"
                f"Total_StrainEnergy: \t{script_e(total_energy, 5)}\t\n")


class BESO:
    """Main.cpp:204-333 for several designs side by side: every iteration
    runs all designs' worlds as one batch of instances on one engine and
    smooths the objectives with mpm_eso_filter.  `simulate(scenes)` (a list
    of (strain energy, maximum-history strain energy) per scene) and
    `smooth` (see Canvas2D.filterObjective_Smooth) replace the GPU steps."""

    def __init__(self, designs, order="pair", scene_kw=None, device=0, use_graph=True, simulate=None, smooth=None):
        self.designs = list(designs)
        self.simulate = simulate
        self.smooth = smooth
        self.order = order
        self.scene_kw = dict(scene_kw or {})
        self.device = device
        self.use_graph = use_graph
        self.iteration = 0
        self.last_scenes = None
        self.timing = {}

    def _simulate(self, scs):
        """Each world run to its end time in requests of half the console
        interval, as GraphicsEngine::gameLoop drives runSimulation
        (GraphicsEngine_gameLoop.cpp:30-52).  Returns per design (Ws, Wmax)."""
        import time as _time
        t0 = _time.perf_counter()
        out = []
        if self.simulate is not None:
            out = self.simulate(scs)
        else:
            from .engine import Engine
            eng = Engine(device=self.device, order=self.order)
            eng.load_batch(scs)   # one instance per design, also for one design (graph replay)
            request = 0.5 * scs[0].console_interval
            while eng.run_batch(request, use_graph=self.use_graph) == 0:
                pass
            eng.synchronize()
            Ws, Wmax = eng.get_particle("ENERGY_STRAIN")[0], eng.get_particle("ENERGY_STRAIN_MAX")[0]
            o = 0
            for sc in scs:
                out.append((Ws[o:o + sc.n].copy(), Wmax[o:o + sc.n].copy()))
                o += sc.n
            self.timing["steps"] = eng.health()["steps"]   # steps of the batch: every design took each
            eng.close()
        self.timing["simulate_s"] = _time.perf_counter() - t0
        return out

    def step(self):
        """One global iteration for every design; returns the log lines."""
        import time as _time
        scs = [eso_scene(d.canvas, **self.scene_kw) for d in self.designs]
        self.last_scenes = scs
        res = self._simulate(scs)
        lines = []
        t_filter = 0.0
        for d, sc, (Ws, Wmax) in zip(self.designs, scs, res):
            total = float(np.cumsum(Ws)[-1]) if Ws.size else 0.0   # Main.cpp:238-245, in particle order
            d.fold_objective(sc.info["voxel_ids"], Wmax)
            t1 = _time.perf_counter()
            d.canvas.filterObjective_Smooth(d.radius, device=self.device, smooth=self.smooth)
            t_filter += _time.perf_counter() - t1
            d.update()
            line = d.log_line(self.iteration, total)
            d.log.append(line)
            lines.append(line)
        self.timing["filter_s"] = t_filter
        self.iteration += 1
        return lines
