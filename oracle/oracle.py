"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker or the timed CPU baseline.  The
product path never imports it.  Method names match code_mpm_amd.Engine so a
Scene applies to either.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")

from code_mpm_amd.capi import MON, NF, PF, STRAIN, Health, Material, Monitors, make_material  # noqa: E402

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    lib = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    I32, I64, DBL = ctypes.c_int32, ctypes.c_int64, ctypes.c_double
    sig = {
        "oracle_create": [ctypes.POINTER(P)],
        "oracle_destroy": [P],
        "oracle_set_grid": [P, P, P],
        "oracle_set_node_boundary": [P, P, P, I64],
        "oracle_set_materials": [P, P, I32],
        "oracle_set_particle_count": [P, I64],
        "oracle_set_particle_field": [P, I32, P, I64],
        "oracle_get_particle_field": [P, I32, P, I64],
        "oracle_set_particle_material": [P, P, I64],
        "oracle_set_displacement_control": [P, P, P, P, I64],
        "oracle_set_monitor_list": [P, I32, P, I64],
        "oracle_set_timeline": [P, P, P, I32],
        "oracle_set_time": [P, DBL, DBL, DBL],
        "oracle_set_damping": [P, DBL, DBL],
        "oracle_step": [P, DBL, I32, I32],
        "oracle_set_order": [P, I32],
        "oracle_set_variant": [P, I32],
        "oracle_run": [P, DBL, I32, ctypes.POINTER(I32)],
        "oracle_get_time": [P, ctypes.POINTER(DBL), ctypes.POINTER(I64)],
        "oracle_get_node_count": [P, ctypes.POINTER(I64)],
        "oracle_get_node_field": [P, I32, P, I64],
        "oracle_get_monitors": [P, ctypes.POINTER(Monitors)],
        "oracle_get_health": [P, ctypes.POINTER(Health)],
        "oracle_find_adjacent": [P, P, P, P, P],
        "oracle_bases": [P, P, P, P, P],
        "oracle_constitutive_batch": [ctypes.POINTER(Material), P, P, P, P, P, P, I64],
        "oracle_timeline_velocity": [P, P, I32, DBL, P],
        "oracle_set_strain_formulation": [P, I32],
        "oracle_matrix_log": [P, P],
        "oracle_matrix_exp": [P, P],
        "oracle_large_strain_point": [P, P, DBL, P, P],
        "oracle_eso_filter": [P, P, P, I32, I32, DBL, DBL, P],
    }
    for name, args in sig.items():
        fn = getattr(lib, name)
        fn.restype = ctypes.c_int32
        fn.argtypes = args
    _lib = lib
    return lib


VARIANT = {"doublepass": 0, "doublepass_contact": 1, "singlepass": 2}


def _p(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.c_void_p)


def _ok(rc, what):
    if rc != 0:
        raise RuntimeError(f"oracle {what} failed: {rc}")


class OracleEngine:
    """The reference's DoublePass step on the CPU (single thread = the
    reference's summation order; more threads = its OpenMP lock scatter)."""

    def __init__(self, threads=1, order="particle", variant="doublepass"):
        self._lib = load()
        h = ctypes.c_void_p()
        _ok(self._lib.oracle_create(ctypes.byref(h)), "create")
        self._h = h
        _ok(self._lib.oracle_set_order(self._h, {"particle": 0, "cell": 1, "tile": 2, "pair": 3}[order]), "order")
        self.set_variant(variant)
        self.threads = int(threads)
        self.n = 0
        self.nnodes = 0

    def close(self):
        if getattr(self, "_h", None):
            self._lib.oracle_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_grid(self, length, cells):
        L = np.ascontiguousarray(length, np.float64)
        C = np.ascontiguousarray(cells, np.int32)
        _ok(self._lib.oracle_set_grid(self._h, _p(L), _p(C)), "set_grid")
        n = ctypes.c_int64()
        self._lib.oracle_get_node_count(self._h, ctypes.byref(n))
        self.nnodes = n.value

    def set_node_boundary(self, fixed=None, friction=None):
        f = None if fixed is None else np.ascontiguousarray(fixed, np.uint8)
        mu = None if friction is None else np.ascontiguousarray(friction, np.float64)
        _ok(self._lib.oracle_set_node_boundary(self._h, _p(f), _p(mu), self.nnodes), "set_node_boundary")

    def set_materials(self, materials):
        arr = (Material * max(1, len(materials)))(*[make_material(m) for m in materials])
        _ok(self._lib.oracle_set_materials(self._h, ctypes.cast(arr, ctypes.c_void_p), len(materials)), "materials")

    def set_particle_count(self, n):
        _ok(self._lib.oracle_set_particle_count(self._h, int(n)), "particle_count")
        self.n = int(n)

    def set_particle(self, name, values):
        fid, nc = PF[name]
        a = np.ascontiguousarray(np.asarray(values, np.float64).reshape(nc, self.n))
        _ok(self._lib.oracle_set_particle_field(self._h, fid, _p(a), self.n), f"set {name}")

    def set_particle_material(self, mat):
        a = np.ascontiguousarray(mat, np.int32)
        _ok(self._lib.oracle_set_particle_material(self._h, _p(a), self.n), "material")

    def set_displacement_control(self, index, axes, multiplier):
        i = np.ascontiguousarray(index, np.int64)
        a = np.ascontiguousarray(axes, np.uint8)
        m = np.ascontiguousarray(multiplier, np.float32)
        _ok(self._lib.oracle_set_displacement_control(self._h, _p(i), _p(a), _p(m), len(i)), "dc")

    def set_monitor_list(self, kind, index):
        i = np.ascontiguousarray(index, np.int64)
        _ok(self._lib.oracle_set_monitor_list(self._h, MON[kind], _p(i), len(i)), "monitor")

    def set_timeline(self, t, v):
        t = np.ascontiguousarray(t, np.float64)
        v = np.ascontiguousarray(v, np.float64).reshape(-1, 3)
        _ok(self._lib.oracle_set_timeline(self._h, _p(t), _p(v), len(t)), "timeline")

    def set_time(self, time, dt_max, time_end):
        _ok(self._lib.oracle_set_time(self._h, float(time), float(dt_max), float(time_end)), "time")

    def set_damping(self, damping, mass_min=0.0):
        _ok(self._lib.oracle_set_damping(self._h, float(damping), float(mass_min)), "damping")

    def load_scene(self, scene):
        scene.apply(self)
        self.scene = scene

    def set_variant(self, variant):
        """doublepass | doublepass_contact | singlepass: the reference's three step loops."""
        _ok(self._lib.oracle_set_variant(self._h, VARIANT[variant]), "variant")
        self.variant = variant

    def set_strain_formulation(self, formulation):
        """small | large: runSimulation_G2P_P2_SmallStrain or _LargeStrain."""
        _ok(self._lib.oracle_set_strain_formulation(self._h, STRAIN[formulation]), "strain formulation")
        self.strain = formulation

    def step(self, dt, nsteps=1):
        _ok(self._lib.oracle_step(self._h, float(dt), int(nsteps), self.threads), "step")

    def run(self, dt_total):
        fin = ctypes.c_int32()
        _ok(self._lib.oracle_run(self._h, float(dt_total), self.threads, ctypes.byref(fin)), "run")
        return fin.value

    def getTime_Current(self):
        t = ctypes.c_double()
        c = ctypes.c_int64()
        self._lib.oracle_get_time(self._h, ctypes.byref(t), ctypes.byref(c))
        return t.value

    def get_particle(self, name):
        fid, nc = PF[name]
        out = np.empty((nc, self.n))
        _ok(self._lib.oracle_get_particle_field(self._h, fid, _p(out), self.n), f"get {name}")
        return out

    def get_node(self, name):
        fid, nc = NF[name]
        out = np.empty((nc, self.nnodes))
        _ok(self._lib.oracle_get_node_field(self._h, fid, _p(out), self.nnodes), f"get node {name}")
        return out

    def monitors(self):
        m = Monitors()
        _ok(self._lib.oracle_get_monitors(self._h, ctypes.byref(m)), "monitors")
        return m.as_dict()

    def health(self):
        h = Health()
        _ok(self._lib.oracle_get_health(self._h, ctypes.byref(h)), "health")
        return h.as_dict()


def find_adjacent(length, cells, x):
    lib = load()
    L = np.ascontiguousarray(length, np.float64)
    C = np.ascontiguousarray(cells, np.int32)
    X = np.ascontiguousarray(x, np.float64)
    out = np.zeros(8, np.int32)
    cnt = np.zeros(1, np.int32)
    _ok(lib.oracle_find_adjacent(_p(L), _p(C), _p(X), _p(out), _p(cnt)), "find_adjacent")
    return out[: cnt[0]].copy()


def bases(h, xp, xn):
    lib = load()
    H = np.ascontiguousarray(h, np.float64)
    A = np.ascontiguousarray(xp, np.float64)
    B = np.ascontiguousarray(xn, np.float64)
    N = np.zeros(1)
    g = np.zeros(3)
    _ok(lib.oracle_bases(_p(H), _p(A), _p(B), _p(N), _p(g)), "bases")
    return N[0], g


def constitutive_batch(material, backstress, stress, dstrain):
    lib = load()
    R = np.ascontiguousarray(backstress, np.float64)
    s = np.ascontiguousarray(stress, np.float64).reshape(-1, 6)
    d = np.ascontiguousarray(dstrain, np.float64).reshape(-1, 6)
    n = s.shape[0]
    ds, dp, dR = np.empty((n, 6)), np.empty((n, 6)), np.empty(n)
    m = make_material(material)
    _ok(lib.oracle_constitutive_batch(ctypes.byref(m), _p(R), _p(s), _p(d), _p(ds), _p(dp), _p(dR), n), "cr")
    return ds, dp, dR


def timeline_velocity(t, v, time):
    lib = load()
    T = np.ascontiguousarray(t, np.float64)
    V = np.ascontiguousarray(v, np.float64).reshape(-1, 3)
    out = np.zeros(3)
    _ok(lib.oracle_timeline_velocity(_p(T), _p(V), len(T), float(time), _p(out)), "timeline")
    return out


def _m3(a):
    """3x3 (row, col) array -> glm column-major 9-vector."""
    return np.ascontiguousarray(np.asarray(a, np.float64).T.reshape(9))


def matrix_log(a):
    """ConstitutiveRelation::getLog of a 3x3 matrix (8-term series)."""
    out = np.empty(9)
    _ok(load().oracle_matrix_log(_p(_m3(a)), _p(out)), "matrix_log")
    return out.reshape(3, 3).T.copy()


def matrix_exp(a):
    """ConstitutiveRelation::getExp of a 3x3 matrix (9-term series)."""
    out = np.empty(9)
    _ok(load().oracle_matrix_exp(_p(_m3(a)), _p(out)), "matrix_exp")
    return out.reshape(3, 3).T.copy()


LARGE_STATE = (("F", 9), ("Fe", 9), ("sig", 6), ("eps", 6), ("epsp", 6), ("R", 1), ("V", 1), ("V0", 1),
               ("Ws", 1), ("Wp", 1))


def large_strain_point(material, L, dt, state):
    """One particle's large-strain update after the node gather.  L is the
    3x3 velocity gradient (row, col); state is a dict with F and Fe (3x3,
    row, col), sig / eps / epsp (6) and R, V, V0, Ws, Wp; material None =
    no material.  Returns (new state, yielded)."""
    flat = []
    for k, nc in LARGE_STATE:
        v = state[k]
        flat.extend(_m3(v) if nc == 9 else np.atleast_1d(np.asarray(v, np.float64)).tolist())
    buf = np.ascontiguousarray(flat, np.float64)
    y = np.zeros(1, np.int32)
    m = None if material is None else make_material(material)
    mp = None if m is None else ctypes.byref(m)
    _ok(load().oracle_large_strain_point(mp, _p(_m3(L)), float(dt), _p(buf), _p(y)), "large_strain_point")
    out, o = {}, 0
    for k, nc in LARGE_STATE:
        v = buf[o:o + nc]
        out[k] = v.reshape(3, 3).T.copy() if nc == 9 else (v.copy() if nc > 1 else float(v[0]))
        o += nc
    return out, bool(y[0])


def eso_filter(active, eso, objective, nx, ny, offset, radius):
    """Canvas2D_CC::filterObjective_Smooth as the reference runs it (all
    pairs of active voxels)."""
    a = np.ascontiguousarray(active, np.uint8)
    e = np.ascontiguousarray(eso, np.uint8)
    o = np.ascontiguousarray(objective, np.float64)
    out = np.empty(int(nx) * int(ny))
    _ok(load().oracle_eso_filter(_p(a), _p(e), _p(o), int(nx), int(ny), float(offset), float(radius), _p(out)),
        "eso_filter")
    return out
