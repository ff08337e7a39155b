"""ctypes binding of the C ABI in include/mpm.h (libmpm_amd.so).

Loading fails loudly when the library has not been built: there is no CPU
fallback for the timestep.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmpm_amd.so")

MPM_OK = 0
ERRORS = {-1: "MPM_ERR_ARG", -2: "MPM_ERR_HIP", -3: "MPM_ERR_STATE", -4: "MPM_ERR_NOMEM", -5: "MPM_ERR_RANGE"}

# material types (Material_BC.h:6-11)
ELASTIC, NEOHOOKEAN, VISCOELASTIC, PLASTIC, VONMISESHARDENING, GASS = range(6)

# particle fields and their component counts
PF = dict(
    POSITION=(0, 3), VELOCITY=(1, 3), FORCE_EXTERNAL=(2, 3), MASS=(3, 1), VOLUME=(4, 1), VOLUME_INITIAL=(5, 1),
    DEFORMATION_GRADIENT=(6, 9), STRAIN=(7, 6), STRESS=(8, 6), STRAIN_PLASTIC=(9, 6), BACKSTRESS=(10, 1),
    ENERGY_STRAIN=(11, 1), ENERGY_STRAIN_MAX=(12, 1), ENERGY_STRAIN_TI=(13, 1), ENERGY_PLASTIC=(14, 1),
    DEFORMATION_GRADIENT_ELASTIC=(15, 9),
)
NF = dict(MASS=(0, 1), MOMENTUM=(1, 3), VELOCITY=(2, 3), FORCE=(3, 3), FORCE_TEMP=(4, 3), ACTIVE=(5, 1))
MON = dict(DISPLACEMENT=0, MOMENTUM=1, ENERGY=2, STRESS=3)
VARIANT = dict(doublepass=0, doublepass_contact=1, singlepass=2)
STRAIN = dict(small=0, large=1)
ORDER = dict(particle=0, cell=1, tile=2, pair=3)
KSLOTS = ("bin", "p2g_partials", "p2g_nodes", "g2p_velocity", "remap_partials", "remap_nodes", "g2p_kinematics",
          "g2p_constitutive")
PHASES = ("bin", "p2g_grid", "g2p_velocity", "remap_grid", "g2p_strain")


class Material(ctypes.Structure):
    _fields_ = [
        ("type", ctypes.c_int32),
        ("id", ctypes.c_int32),
        ("density", ctypes.c_double),
        ("elastic_modulus", ctypes.c_double),
        ("poisson_ratio", ctypes.c_double),
        ("yield_stress", ctypes.c_double),
        ("viscosity", ctypes.c_double),
        ("hardening_c0", ctypes.c_double),
        ("hardening_c1", ctypes.c_double),
    ]


class Monitors(ctypes.Structure):
    _fields_ = [
        ("time", ctypes.c_double),
        ("position_y", ctypes.c_double),
        ("force_y", ctypes.c_double),
        ("momentum_y", ctypes.c_double),
        ("kinetic_energy", ctypes.c_double),
        ("strain_energy", ctypes.c_double),
        ("plastic_energy", ctypes.c_double),
        ("strain_y", ctypes.c_double),
        ("stress_y", ctypes.c_double),
        ("has_position", ctypes.c_int32),
        ("has_momentum", ctypes.c_int32),
        ("has_energy", ctypes.c_int32),
        ("has_stress", ctypes.c_int32),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class Health(ctypes.Structure):
    _fields_ = [
        ("steps", ctypes.c_int64),
        ("nonfinite_particles", ctypes.c_int64),
        ("outside_grid", ctypes.c_int64),
        ("null_material", ctypes.c_int64),
        ("plastic_points", ctypes.c_int64),
        ("surrogate_calls", ctypes.c_int64),
        ("surrogate_fallbacks", ctypes.c_int64),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


P = ctypes.c_void_p
I32 = ctypes.c_int32
I64 = ctypes.c_int64
DBL = ctypes.c_double
PI32 = ctypes.POINTER(ctypes.c_int32)
PI64 = ctypes.POINTER(ctypes.c_int64)
PDBL = ctypes.POINTER(ctypes.c_double)

# name -> (restype, argtypes); the exported symbol set of include/mpm.h
SIGNATURES = {
    "mpm_create": (I32, [ctypes.POINTER(P), I32]),
    "mpm_destroy": (I32, [P]),
    "mpm_last_error": (ctypes.c_char_p, [P]),
    "mpm_set_grid": (I32, [P, PDBL, PI32]),
    "mpm_set_node_boundary": (I32, [P, P, P, I64]),
    "mpm_set_materials": (I32, [P, P, I32]),
    "mpm_set_particle_count": (I32, [P, I64]),
    "mpm_get_particle_count": (I32, [P, PI64]),
    "mpm_set_particle_field": (I32, [P, I32, P, I64]),
    "mpm_get_particle_field": (I32, [P, I32, P, I64]),
    "mpm_set_particle_material": (I32, [P, P, I64]),
    "mpm_set_displacement_control": (I32, [P, P, P, P, I64]),
    "mpm_set_monitor_list": (I32, [P, I32, P, I64]),
    "mpm_set_timeline": (I32, [P, P, P, I32]),
    "mpm_set_time": (I32, [P, DBL, DBL, DBL]),
    "mpm_set_damping": (I32, [P, DBL, DBL]),
    "mpm_step": (I32, [P, DBL, I32]),
    "mpm_step_graph": (I32, [P, I32]),
    "mpm_set_summation_order": (I32, [P, I32]),
    "mpm_set_variant": (I32, [P, I32]),
    "mpm_get_variant": (I32, [P, PI32]),
    "mpm_set_strain_formulation": (I32, [P, I32]),
    "mpm_eso_filter": (I32, [I32, P, P, P, I32, I32, DBL, DBL, P]),
    "mpm_get_strain_formulation": (I32, [P, PI32]),
    "mpm_set_tile_shape": (I32, [P, PI32]),
    "mpm_get_tile_shape": (I32, [P, PI32]),
    "mpm_autotune": (I32, [P, I32, PI32, PDBL]),
    "mpm_set_instances": (I32, [P, I32, P, P, P]),
    "mpm_get_instance_layout": (I32, [P, I32, PI64, PI64, PI64]),
    "mpm_set_instance_timeline": (I32, [P, I32, P, P, I32]),
    "mpm_set_instance_time": (I32, [P, I32, DBL, DBL, DBL]),
    "mpm_get_instance_time": (I32, [P, I32, PDBL]),
    "mpm_get_instance_monitors": (I32, [P, I32, ctypes.POINTER(Monitors)]),
    "mpm_run": (I32, [P, DBL, PI32]),
    "mpm_step_timed": (I32, [P, DBL, I32, PDBL]),
    "mpm_step_phase": (I32, [P, I32, DBL]),
    "mpm_step_timed_kernels": (I32, [P, DBL, I32, PDBL]),
    "mpm_synchronize": (I32, [P]),
    "mpm_get_time": (I32, [P, PDBL, PI64]),
    "mpm_get_node_count": (I32, [P, PI64]),
    "mpm_get_node_field": (I32, [P, I32, P, I64]),
    "mpm_get_monitors": (I32, [P, ctypes.POINTER(Monitors)]),
    "mpm_get_health": (I32, [P, ctypes.POINTER(Health)]),
    "mpm_reset_health": (I32, [P]),
    "mpm_constitutive_batch": (I32, [I32, ctypes.POINTER(Material), P, P, P, P, P, P, I64]),
    "mpm_surrogate_param_count": (I64, []),
    "mpm_surrogate_create": (I32, [ctypes.POINTER(P), I32, ctypes.c_uint64]),
    "mpm_surrogate_destroy": (I32, [P]),
    "mpm_surrogate_set_params": (I32, [P, P, I64]),
    "mpm_surrogate_get_params": (I32, [P, P, I64]),
    "mpm_surrogate_set_normalization": (I32, [P, DBL, DBL, P, P, P, P, P, P]),
    "mpm_surrogate_forward": (I32, [P, P, P, I64]),
    "mpm_surrogate_gradient": (I32, [P, P, P, I64, P, P]),
    "mpm_surrogate_gradient_noalloc": (I32, [P, P, P, I64, P, P]),
    "mpm_surrogate_load_training_set": (I32, [P, P, P, I64]),
    "mpm_surrogate_train": (I32, [P, I32, I32, ctypes.c_float, ctypes.c_uint64, P]),
    "mpm_set_surrogate": (I32, [P, P, I32, DBL]),
    "mpm_set_sample_capture": (I32, [P, I64]),
    "mpm_get_captured_samples": (I32, [P, P, I64, PI64]),
}

_lib = None


def load(path=LIB_PATH):
    """Load libmpm_amd.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is missing: build the HIP extension first (python -m code_mpm_amd.build); "
            "the timestep has no CPU fallback")
    lib = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class MPMError(RuntimeError):
    pass


def check(rc, handle=None, what=""):
    if rc != MPM_OK:
        msg = load().mpm_last_error(handle).decode(errors="replace")
        raise MPMError(f"{what}: {ERRORS.get(rc, rc)}: {msg}")


def ptr(a):
    """Address of a C-contiguous numpy array (None for None)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "array must be C-contiguous"
    return a.ctypes.data_as(ctypes.c_void_p)


def make_material(m):
    """Material struct from a dict with the Material_BC field names."""
    return Material(int(m.get("type", 0)), int(m.get("id", 0)), float(m.get("density", 0.0)),
                    float(m.get("elastic_modulus", 0.0)), float(m.get("poisson_ratio", 0.0)),
                    float(m.get("yield_stress", 0.0)), float(m.get("viscosity", 0.0)),
                    float(m.get("hardening_c0", 0.0)), float(m.get("hardening_c1", 0.0)))


def constitutive_batch(material, backstress, stress, dstrain, device=0):
    """Run the device return map on n points; returns (dstress, dstrain_plastic, dbackstress)."""
    lib = load()
    R = np.ascontiguousarray(backstress, dtype=np.float64)
    s = np.ascontiguousarray(stress, dtype=np.float64).reshape(-1, 6)
    d = np.ascontiguousarray(dstrain, dtype=np.float64).reshape(-1, 6)
    n = s.shape[0]
    assert R.shape == (n,) and d.shape == (n, 6)
    ds = np.empty((n, 6))
    dp = np.empty((n, 6))
    dR = np.empty(n)
    mat = make_material(material)
    rc = lib.mpm_constitutive_batch(device, ctypes.byref(mat), ptr(R), ptr(s), ptr(d), ptr(ds), ptr(dp), ptr(dR), n)
    check(rc, None, "mpm_constitutive_batch")
    return ds, dp, dR


def eso_filter(active, eso, objective, nx, ny, offset, radius, device=0):
    """Canvas2D_CC::filterObjective_Smooth on the GPU; arrays of nx * ny
    voxels (index x + y * nx).  Returns the new objective array."""
    lib = load()
    a = np.ascontiguousarray(active, dtype=np.uint8)
    e = np.ascontiguousarray(eso, dtype=np.uint8)
    o = np.ascontiguousarray(objective, dtype=np.float64)
    assert a.size == e.size == o.size == nx * ny
    out = np.empty(nx * ny)
    rc = lib.mpm_eso_filter(int(device), ptr(a), ptr(e), ptr(o), int(nx), int(ny), float(offset), float(radius),
                            ptr(out))
    check(rc, None, "mpm_eso_filter")
    return out
