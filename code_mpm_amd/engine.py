"""Python host API over the C ABI: the PhysicsEngine of the reference
(PhysicsEngine.h:30-122) with the state on the GPU.

``Engine`` mirrors the reference's member functions that drive the step:

- ``runSimulation_Classic_DoublePass_MPLocks(dt_total)`` returns 0 while
  ``time < time_end`` and 1 when done (PhysicsEngine_runSimulation_Classic_
  DoublePass_MPLocks.cpp:806-814) and calls ``reportConsole`` when the console
  interval has elapsed (same file :796-800);
- ``getTime_*`` / ``getCount_*`` getters (PhysicsEngine.h:66-75);
- ``get_particle`` / ``get_node`` are the lazy device->host mirrors behind
  ``getMaterialPoints`` / ``getGridPoints`` (PhysicsEngine.h:76-78).

Scenes (``code_mpm_amd.scenes.Scene``) play the part of the reference's
``initializeWorld_*`` functions.
"""
import ctypes
import time

import numpy as np

from . import capi
from .capi import MON, NF, PF, check, ptr


def script(x, precision):
    """Script(double, precision) of Definitions.h:46-60: scientific notation,
    that many digits after the point, explicit sign."""
    return f"{x:+.{precision}e}"


def console_line(m, runtime, description=""):
    """The line PhysicsEngine::reportConsole writes (PhysicsEngine.cpp:42-151)
    for the monitors m: the description, then tab-separated columns for time
    and runtime, Position_y (displacement monitor set), Force_y, momentum_y,
    Strain_y / Stress_y and the three energies, and a newline.  The
    principal-value columns are not produced."""
    s = description + "\ttime: " + script(m["time"], 4) + "\tRuntime: " + script(runtime, 3)
    if m["has_position"]:
        s += "\tPosition_y: " + script(m["position_y"], 14)
    s += "\tForce_y: " + script(m["force_y"], 4)
    if m["has_momentum"]:
        s += "\t momentum_y: " + script(m["momentum_y"], 4)
    if m["has_stress"]:
        s += "\tStrain_y: " + script(m["strain_y"], 4) + "\tStress_y: " + script(m["stress_y"], 4)
    if m["has_energy"]:
        s += ("\tKinetic Energy: " + script(m["kinetic_energy"], 4) + "\tStrain Energy: " +
              script(m["strain_energy"], 4) + "\tPlastic Energy: " + script(m["plastic_energy"], 4))
    return s + "\n"


class Engine:
    """One simulation on one GPU (one HIP stream)."""

    def __init__(self, device=0, order="pair", variant="doublepass"):
        self._lib = capi.load()
        h = ctypes.c_void_p()
        rc = self._lib.mpm_create(ctypes.byref(h), int(device))
        check(rc, None, "mpm_create")
        self._h = h
        self.set_order(order)
        self.variant = None
        self.set_variant(variant)
        self.n = 0
        self.nnodes = 0
        self.console_interval = None
        self.console_last = -1.0e12
        self.console_log = []
        self.runtime_total = 0.0   # d_Runtime_Total: wall seconds of the loops' steps
        self.console_print = False
        self.log_file = None
        self.scene = None

    def set_order(self, order):
        """Node summation order: "particle" (the reference's one-thread order,
        bit for bit), or a grouped order of per-cell partial sums in a fixed
        association: "pair" (default; the two cells along x paired first),
        "cell" or "tile" (the same terms, rounding differently; fast)."""
        self._check(self._lib.mpm_set_summation_order(self._h, capi.ORDER[order]), "set_order")
        self.order = order

    def set_variant(self, variant):
        """Step loop: "doublepass" (runSimulation_Classic_DoublePass_MPLocks),
        "doublepass_contact" (..._DoublePass_MPLocks_Contact) or "singlepass"
        (..._SinglePass_MP)."""
        if variant != self.variant:
            self._check(self._lib.mpm_set_variant(self._h, capi.VARIANT[variant]), "set_variant")
            self.variant = variant

    def set_strain_formulation(self, formulation):
        """G2P pass 2 of the DoublePass loop: "small"
        (runSimulation_G2P_P2_SmallStrain, the reference's) or "large"
        (runSimulation_G2P_P2_LargeStrain, logarithmic strains; DoublePass
        loop and exact return map only)."""
        self._check(self._lib.mpm_set_strain_formulation(self._h, capi.STRAIN[formulation]),
                    "set_strain_formulation")
        self.strain = formulation

    # -- lifetime ---------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            self._lib.mpm_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc, what):
        check(rc, self._h, what)

    # -- set-up -------------------------------------------------------------
    def set_grid(self, length, cells):
        L = np.ascontiguousarray(length, dtype=np.float64)
        C = np.ascontiguousarray(cells, dtype=np.int32)
        self._check(self._lib.mpm_set_grid(self._h, L.ctypes.data_as(capi.PDBL), C.ctypes.data_as(capi.PI32)),
                    "set_grid")
        n = ctypes.c_int64()
        self._lib.mpm_get_node_count(self._h, ctypes.byref(n))
        self.nnodes = n.value

    def set_node_boundary(self, fixed=None, friction=None):
        f = None if fixed is None else np.ascontiguousarray(fixed, dtype=np.uint8)
        mu = None if friction is None else np.ascontiguousarray(friction, dtype=np.float64)
        self._check(self._lib.mpm_set_node_boundary(self._h, ptr(f), ptr(mu), self.nnodes), "set_node_boundary")

    def set_materials(self, materials):
        arr = (capi.Material * max(1, len(materials)))(*[capi.make_material(m) for m in materials])
        self._check(self._lib.mpm_set_materials(self._h, ctypes.cast(arr, ctypes.c_void_p), len(materials)),
                    "set_materials")

    def set_particle_count(self, n):
        self._check(self._lib.mpm_set_particle_count(self._h, int(n)), "set_particle_count")
        self.n = int(n)

    def set_particle(self, name, values):
        fid, nc = PF[name]
        a = np.ascontiguousarray(np.asarray(values, dtype=np.float64).reshape(nc, self.n))
        self._check(self._lib.mpm_set_particle_field(self._h, fid, ptr(a), self.n), f"set_particle {name}")

    def set_particle_material(self, mat):
        a = np.ascontiguousarray(mat, dtype=np.int32)
        self._check(self._lib.mpm_set_particle_material(self._h, ptr(a), self.n), "set_particle_material")

    def set_displacement_control(self, index, axes, multiplier):
        i = np.ascontiguousarray(index, dtype=np.int64)
        a = np.ascontiguousarray(axes, dtype=np.uint8)
        m = np.ascontiguousarray(multiplier, dtype=np.float32)
        self._check(self._lib.mpm_set_displacement_control(self._h, ptr(i), ptr(a), ptr(m), len(i)),
                    "set_displacement_control")

    def set_monitor_list(self, kind, index):
        i = np.ascontiguousarray(index, dtype=np.int64)
        self._check(self._lib.mpm_set_monitor_list(self._h, MON[kind], ptr(i), len(i)), "set_monitor_list")

    def set_timeline(self, t, v):
        t = np.ascontiguousarray(t, dtype=np.float64)
        v = np.ascontiguousarray(v, dtype=np.float64).reshape(-1, 3)
        self._check(self._lib.mpm_set_timeline(self._h, ptr(t), ptr(v), len(t)), "set_timeline")

    def set_time(self, time, dt_max, time_end):
        self._check(self._lib.mpm_set_time(self._h, float(time), float(dt_max), float(time_end)), "set_time")

    def set_damping(self, damping, mass_min=0.0):
        self._check(self._lib.mpm_set_damping(self._h, float(damping), float(mass_min)), "set_damping")

    def load_scene(self, scene):
        """Upload a Scene (the equivalent of an initializeWorld_* call)."""
        scene.apply(self)
        self.scene = scene
        self.console_interval = scene.console_interval
        self.console_last = 0.0

    def load_batch(self, scenes_):
        """Upload several independent scenes as one batch: one set of launches
        per step for all of them (the BESO driver's stream of independent
        worlds, Main.cpp:209-336, run side by side).  Materials equal by value
        are shared; control and monitor lists are offset to global indices."""
        scenes_ = list(scenes_)
        L = np.ascontiguousarray([sc.length for sc in scenes_], np.float64).ravel()
        C = np.ascontiguousarray([sc.cells for sc in scenes_], np.int32).ravel()
        N = np.ascontiguousarray([sc.n for sc in scenes_], np.int64)
        self._check(self._lib.mpm_set_instances(self._h, len(scenes_), ptr(L), ptr(C), ptr(N)), "set_instances")
        n = ctypes.c_int64()
        self._lib.mpm_get_node_count(self._h, ctypes.byref(n))
        self.nnodes = n.value
        self.n = int(N.sum())
        mats, key_of = [], {}
        mat_ids = []
        for sc in scenes_:
            remap = []
            for m in sc.materials:
                key = tuple(sorted((k, float(v)) for k, v in m.items()))
                if key not in key_of:
                    key_of[key] = len(mats)
                    mats.append(m)
                remap.append(key_of[key])
            remap = np.asarray(remap + [-1], np.int32)
            mat_ids.append(np.where(sc.material >= 0, remap[np.clip(sc.material, 0, len(remap) - 2)], -1))
        self.set_materials(mats)
        self.set_node_boundary(np.concatenate([sc.fixed for sc in scenes_]),
                               np.concatenate([sc.friction for sc in scenes_]))
        names = set()
        for sc in scenes_:
            names |= set(sc.fields)
        from .capi import PF as _PF
        for name in names:
            nc = _PF[name][1]
            parts = []
            for sc in scenes_:
                if name in sc.fields:
                    parts.append(np.asarray(sc.fields[name], np.float64).reshape(nc, sc.n))
                else:
                    d = np.zeros((nc, sc.n))
                    if name == "DEFORMATION_GRADIENT":
                        d[[0, 4, 8]] = 1.0
                    parts.append(d)
            self.set_particle(name, np.concatenate(parts, axis=1))
        self.set_particle_material(np.concatenate(mat_ids).astype(np.int32))
        offs = np.concatenate([[0], np.cumsum(N)[:-1]])
        self.set_displacement_control(np.concatenate([sc.dc_index + o for sc, o in zip(scenes_, offs)]),
                                      np.concatenate([sc.dc_axes for sc in scenes_]),
                                      np.concatenate([sc.dc_mult for sc in scenes_]))
        kinds = set()
        for sc in scenes_:
            kinds |= set(sc.monitors)
        for kind in kinds:
            self.set_monitor_list(kind, np.concatenate(
                [np.asarray(sc.monitors.get(kind, []), np.int64) + o for sc, o in zip(scenes_, offs)]))
        for k, sc in enumerate(scenes_):
            t = np.ascontiguousarray(sc.timeline_t, np.float64)
            v = np.ascontiguousarray(sc.timeline_v, np.float64).reshape(-1, 3)
            self._check(self._lib.mpm_set_instance_timeline(self._h, k, ptr(t), ptr(v), len(t)), "instance timeline")
            self._check(self._lib.mpm_set_instance_time(self._h, k, float(sc.time0), float(sc.dt), float(sc.time_end)),
                        "instance time")
        self.set_damping(scenes_[0].damping, scenes_[0].mass_min)
        self.batch = scenes_
        self.scene = None
        return offs

    def instance_monitors(self, k):
        m = capi.Monitors()
        self._check(self._lib.mpm_get_instance_monitors(self._h, int(k), ctypes.byref(m)), "instance monitors")
        return m.as_dict()

    def instance_time(self, k):
        t = ctypes.c_double()
        self._check(self._lib.mpm_get_instance_time(self._h, int(k), ctypes.byref(t)), "instance time")
        return t.value

    def set_tile_shape(self, shape):
        a = np.ascontiguousarray(shape, np.int32)
        self._check(self._lib.mpm_set_tile_shape(self._h, a.ctypes.data_as(capi.PI32)), "set_tile_shape")

    def tile_shape(self):
        a = np.zeros(3, np.int32)
        self._check(self._lib.mpm_get_tile_shape(self._h, a.ctypes.data_as(capi.PI32)), "get_tile_shape")
        return tuple(int(v) for v in a)

    def autotune(self, reps=3):
        """One dt_max step that also picks the fastest node tile shape."""
        a = np.zeros(3, np.int32)
        ms = ctypes.c_double()
        self._check(self._lib.mpm_autotune(self._h, int(reps), a.ctypes.data_as(capi.PI32), ctypes.byref(ms)),
                    "autotune")
        return tuple(int(v) for v in a), ms.value

    def step_graph(self, nsteps):
        """nsteps steps replayed from a HIP graph (recorded on first use)."""
        self._check(self._lib.mpm_step_graph(self._h, int(nsteps)), "step_graph")

    # -- the step -------------------------------------------------------------
    def step(self, dt, nsteps=1):
        self._check(self._lib.mpm_step(self._h, float(dt), int(nsteps)), "step")

    def step_phase(self, phase, dt=0.0):
        """One phase of a step taken phase by phase ("bin" opens the step and
        takes its dt; the others follow in the loop's order, and "g2p_strain"
        closes it)."""
        self._check(self._lib.mpm_step_phase(self._h, capi.PHASES.index(phase), float(dt)), f"step_phase {phase}")

    def phase_order(self):
        """The phases of one step in the current loop's order."""
        mid = ("remap_grid", "g2p_velocity") if self.variant == "singlepass" else ("g2p_velocity", "remap_grid")
        return ("bin", "p2g_grid") + mid + ("g2p_strain",)

    def step_timed(self, dt, nsteps=1):
        """Steps with HIP-event timing per phase; returns {phase: total ms}."""
        out = np.zeros(len(capi.PHASES))
        self._check(self._lib.mpm_step_timed(self._h, float(dt), int(nsteps), out.ctypes.data_as(capi.PDBL)),
                    "step_timed")
        return dict(zip(capi.PHASES, out.tolist()))

    def step_timed_kernels(self, dt, nsteps=1):
        """Steps with HIP-event timing per kernel; returns {slot: total ms}."""
        out = np.zeros(len(capi.KSLOTS))
        self._check(self._lib.mpm_step_timed_kernels(self._h, float(dt), int(nsteps),
                                                     out.ctypes.data_as(capi.PDBL)), "step_timed_kernels")
        return dict(zip(capi.KSLOTS, out.tolist()))

    def run(self, dt_total):
        fin = ctypes.c_int32()
        self._check(self._lib.mpm_run(self._h, float(dt_total), ctypes.byref(fin)), "run")
        return fin.value

    def run_batch(self, dt_total, use_graph=False):
        """runSimulation_Classic_DoublePass_MPLocks(dt_total)
        (PhysicsEngine_runSimulation_Classic_DoublePass_MPLocks.cpp:450-458,
        802-814) for every instance of a batch whose instances share one
        d_TimeIncrement_Maximum: steps of dt_max with the last one clamped
        to the remainder, the request accumulated in floating point as the
        reference does.  Runs of full steps go to the device in one call
        (replayed from a HIP graph with use_graph); a clamped step sets every
        instance's step size for that one step.  Returns 0 while any
        instance's time is below its end time, else 1."""
        if not getattr(self, "batch", None):
            raise RuntimeError("run_batch needs a batch (load_batch)")
        dts = {float(sc.dt) for sc in self.batch}
        if len(dts) != 1:
            raise ValueError("run_batch needs one dt_max for every instance")
        dt_max = dts.pop()
        acc, full = 0.0, 0

        def flush():
            nonlocal full
            if full:
                (self.step_graph if use_graph else (lambda k: self.step(0.0, k)))(full)
            full = 0

        while acc < dt_total:
            dt = dt_max
            if dt_max > (dt_total - acc):
                dt = dt_total - acc
            acc += dt
            if dt == dt_max:
                full += 1
                continue
            flush()
            for k, sc in enumerate(self.batch):
                self._check(self._lib.mpm_set_instance_time(self._h, k, self.instance_time(k), dt, float(sc.time_end)),
                            "instance time")
            self.step(0.0, 1)
            for k, sc in enumerate(self.batch):
                self._check(self._lib.mpm_set_instance_time(self._h, k, self.instance_time(k), dt_max,
                                                            float(sc.time_end)), "instance time")
        flush()
        return 0 if any(self.instance_time(k) < sc.time_end for k, sc in enumerate(self.batch)) else 1

    def synchronize(self):
        self._check(self._lib.mpm_synchronize(self._h), "synchronize")

    def runSimulation_Classic_DoublePass_MPLocks(self, dt_total):
        """The reference's step loop with its console report
        (PhysicsEngine_runSimulation_Classic_DoublePass_MPLocks.cpp:439-815)."""
        self.set_variant("doublepass")
        return self._run_loop(dt_total)

    def runSimulation_Classic_DoublePass_MPLocks_Contact(self, dt_total):
        """PhysicsEngine_runSimulation_Classic_DoublePass_MPLocks_Contact.cpp:4-566."""
        self.set_variant("doublepass_contact")
        return self._run_loop(dt_total)

    def runSimulation_Classic_SinglePass_MP(self, dt_total):
        """PhysicsEngine_runSimulation_Classic_SinglePass_MP.cpp:4-404."""
        self.set_variant("singlepass")
        return self._run_loop(dt_total)

    def _run_loop(self, dt_total):
        """Loop control shared by the three (DP:439-458, 790-814; DPC:15-24,
        545-565; SP:16-24, 383-403): dt clamped to the request, console
        report when the interval has passed, 0 while d_Time < d_TimeEnd."""
        acc = 0.0
        dt_max = self.scene.dt if self.scene is not None else None
        time_end = self.scene.time_end if self.scene is not None else float("inf")
        mark = time.perf_counter()
        while acc < dt_total:
            dt = dt_max
            if dt_max > (dt_total - acc):
                dt = dt_total - acc
            acc += dt
            t_before = self.getTime_Current()
            # the console check of the reference sees d_Time before the increment
            self.step(dt, 1)
            if self.console_interval is not None and t_before - self.console_last > self.console_interval:
                self.synchronize()
                self.runtime_total += time.perf_counter() - mark
                self.console_last = t_before
                self.reportConsole(time_override=t_before)
                mark = time.perf_counter()
        self.synchronize()
        self.runtime_total += time.perf_counter() - mark
        if self.getTime_Current() < time_end:
            return 0
        self.console_last = self.getTime_Current()
        self.reportConsole()
        return 1

    def reportConsole(self, time_override=None, description=""):
        """PhysicsEngine::reportConsole (PhysicsEngine.cpp:40-183): the
        monitors, kept in console_log with the line the reference writes
        (console_line); the line is printed when console_print is set and
        appended to log_file when that is set."""
        m = self.monitors()
        if time_override is not None:
            m["time"] = time_override
        m["line"] = console_line(m, self.runtime_total, description)
        self.console_log.append(m)
        if getattr(self, "console_print", False):
            print(m["line"], end="", flush=True)
        if getattr(self, "log_file", None):
            with open(self.log_file, "a") as f:
                f.write(m["line"])
        return m

    # -- getters -------------------------------------------------------------
    def getTime_Current(self):
        t = ctypes.c_double()
        c = ctypes.c_int64()
        self._lib.mpm_get_time(self._h, ctypes.byref(t), ctypes.byref(c))
        return t.value

    def getTime_Cycle(self):
        t = ctypes.c_double()
        c = ctypes.c_int64()
        self._lib.mpm_get_time(self._h, ctypes.byref(t), ctypes.byref(c))
        return c.value

    def getCount_MaterialPoint(self):
        return self.n

    def getCount_GridPoint(self):
        return self.nnodes

    def get_particle(self, name):
        fid, nc = PF[name]
        out = np.empty((nc, self.n))
        self._check(self._lib.mpm_get_particle_field(self._h, fid, ptr(out), self.n), f"get_particle {name}")
        return out

    def get_node(self, name):
        fid, nc = NF[name]
        out = np.empty((nc, self.nnodes))
        self._check(self._lib.mpm_get_node_field(self._h, fid, ptr(out), self.nnodes), f"get_node {name}")
        return out

    def monitors(self):
        m = capi.Monitors()
        self._check(self._lib.mpm_get_monitors(self._h, ctypes.byref(m)), "monitors")
        return m.as_dict()

    def health(self):
        h = capi.Health()
        self._check(self._lib.mpm_get_health(self._h, ctypes.byref(h)), "health")
        return h.as_dict()

    def reset_health(self):
        self._check(self._lib.mpm_reset_health(self._h), "reset_health")

    # -- learned return map ----------------------------------------------------
    def set_surrogate(self, surrogate, material=0, residual_tolerance=1.0e-4):
        """Route `material`'s return map through a trained Surrogate (None: exact map)."""
        h = surrogate.handle if surrogate is not None else None
        self._check(self._lib.mpm_set_surrogate(self._h, h, int(material), float(residual_tolerance)),
                    "set_surrogate")
        self._surrogate = surrogate

    def set_sample_capture(self, capacity):
        self._check(self._lib.mpm_set_sample_capture(self._h, int(capacity)), "set_sample_capture")
        self._capture_capacity = int(capacity)

    def captured_samples_count(self):
        """Rows the capture has produced so far (may exceed its capacity)."""
        cnt = ctypes.c_int64()
        self._check(self._lib.mpm_get_captured_samples(self._h, None, 0, ctypes.byref(cnt)), "captured count")
        return cnt.value

    def captured_samples(self):
        """(rows, produced): rows is (k, 14) = sigma6, dstrain6, R, material."""
        cnt = ctypes.c_int64()
        self._check(self._lib.mpm_get_captured_samples(self._h, None, 0, ctypes.byref(cnt)), "captured count")
        k = min(cnt.value, getattr(self, "_capture_capacity", 0))
        rows = np.empty((k, 14))
        if k:
            self._check(self._lib.mpm_get_captured_samples(self._h, ptr(rows), k, ctypes.byref(cnt)), "captured")
        return rows, cnt.value
