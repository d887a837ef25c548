"""ctypes binding of libsphe.so (include/sphe.h).

The library is built in-tree (make -C sph-erosion_amd) and loaded from
sph-erosion_amd/libsphe.so.  There is no fallback: if the library is missing
or fails to load, every call raises.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
PKG_ROOT = os.path.dirname(HERE)
LIB_PATH = os.path.join(PKG_ROOT, "libsphe.so")

SPHE_OK = 0
ERRORS = {-1: "SPHE_EARG", -2: "SPHE_EHIP", -3: "SPHE_ESTATE", -4: "SPHE_ENOMEM", -5: "SPHE_ERANGE"}

FP32, FP32_NAIVE, FP64, FP32_F64ACC = 0, 1, 2, 3
EROSION_OFF, EROSION_ATOMIC, EROSION_TWO_PHASE = 0, 1, 2
P_MASS, P_VISC, P_SURF_TEN, P_P0, P_GRAV = 0, 1, 2, 3, 4
BUF_POS, BUF_VEL, BUF_DENSITY, BUF_MASS, BUF_PRESS, BUF_SED, BUF_ID = range(7)


class Params(C.Structure):
    _fields_ = [("h", C.c_float), ("mass", C.c_float), ("rest_density", C.c_float),
                ("stiffness", C.c_float), ("visc", C.c_float), ("surf_tens", C.c_float),
                ("gravity", C.c_float * 3), ("dt", C.c_float), ("box_center", C.c_float * 3),
                ("box_half", C.c_float), ("erosion", C.c_int32), ("kc", C.c_float),
                ("ke", C.c_float), ("kd", C.c_float), ("band", C.c_float)]


class TerrainDesc(C.Structure):
    _fields_ = [("nx", C.c_int32), ("nz", C.c_int32), ("origin", C.c_float * 3),
                ("cell", C.c_float), ("heights", C.c_void_p)]


class Particle(C.Structure):
    """FluidParticle (Erosion/fluid_system.h:49-64), 112 bytes."""
    _fields_ = [("Id", C.c_int32), ("Position", C.c_float * 3), ("Velocity", C.c_float * 3),
                ("Acceleration", C.c_float * 3), ("Density", C.c_float), ("Pressure", C.c_float),
                ("PressureForce", C.c_float * 3), ("ViscosityForce", C.c_float * 3),
                ("GravityForce", C.c_float * 3), ("SurfaceForce", C.c_float * 3),
                ("SurfaceNormal", C.c_float * 3), ("NeighbId", C.c_int32)]


assert C.sizeof(Particle) == 112

# name -> (restype, argtypes); every symbol declared in include/sphe.h
P = C.c_void_p
I = C.c_int
SIGNATURES = {
    "sphe_version": (C.c_char_p, []),
    "sphe_params_default": (None, [C.POINTER(Params)]),
    "sphe_create": (I, [I, C.POINTER(P)]),
    "sphe_destroy": (I, [P]),
    "sphe_last_error": (C.c_char_p, [P]),
    "sphe_add_scene": (I, [P, C.POINTER(Params), C.POINTER(TerrainDesc), C.POINTER(I)]),
    "sphe_scene_count": (I, [P, C.POINTER(I)]),
    "sphe_set_params": (I, [P, I, C.POINTER(Params)]),
    "sphe_get_params": (I, [P, I, C.POINTER(Params)]),
    "sphe_param_ptr": (C.POINTER(C.c_float), [P, I, I]),
    "sphe_set_dt": (I, [P, I, C.c_float]),
    "sphe_get_dt": (I, [P, I, C.POINTER(C.c_float)]),
    "sphe_set_origin": (I, [P, I, P]),
    "sphe_get_origin": (I, [P, I, P]),
    "sphe_init_lattice": (I, [P, I, I]),
    "sphe_add_particles": (I, [P, I, I]),
    "sphe_reset": (I, [P, I]),
    "sphe_set_particles": (I, [P, I, I, P, P, P, P]),
    "sphe_set_particles_f64": (I, [P, I, I, P, P, P, P]),
    "sphe_num_particles": (I, [P, I, C.POINTER(I)]),
    "sphe_read_buffer": (I, [P, I, I, P, C.c_size_t]),
    "sphe_buffer": (I, [P, I, I, C.POINTER(P), C.POINTER(C.c_size_t)]),
    "sphe_batch_create": (I, [I, P, P, I, C.POINTER(P)]),
    "sphe_batch_step": (I, [P, I]),
    "sphe_read_state_f64": (I, [P, I, P, P]),
    "sphe_read_sediment_fx": (I, [P, I, P]),
    "sphe_get_particle": (I, [P, I, I, C.POINTER(Particle)]),
    "sphe_set_terrain": (I, [P, I, C.POINTER(TerrainDesc)]),
    "sphe_terrain_load_u8": (I, [P, I, P, I, I, C.c_float]),
    "sphe_terrain_heights": (I, [P, I, P, C.c_size_t]),
    "sphe_terrain_volume_fx": (I, [P, I, P, C.c_size_t]),
    "sphe_terrain_dims": (I, [P, I, C.POINTER(I), C.POINTER(I)]),
    "sphe_fixed_exponent": (I, [P, I, C.POINTER(I)]),
    "sphe_terrain_mesh": (I, [P, I, P, C.POINTER(C.c_size_t), P, C.POINTER(C.c_size_t)]),
    "sphe_terrain_mesh_device": (I, [P, I, C.POINTER(C.c_void_p), C.POINTER(C.c_size_t)]),
    "sphe_positions_device": (I, [P, I, C.POINTER(C.c_void_p), C.POINTER(C.c_size_t)]),
    "sphe_step": (I, [P, I]),
    "sphe_sync": (I, [P]),
    "sphe_use_graph": (I, [P, I]),
    "sphe_mass": (I, [P, I, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_uint64)]),
    "sphe_mass_all": (I, [P, P, P, P, I]),
    "sphe_mass_all_noalloc": (I, [P, P, P, P, I]),
    "sphe_status": (I, [P, C.POINTER(C.c_uint32)]),
    "sphe_profile": (I, [P, I]),
    "sphe_profile_read": (I, [P, C.c_char_p, C.POINTER(C.c_double), C.POINTER(I)]),
    "sphe_profile_names": (I, [P, C.c_char_p, C.c_size_t]),
    "sphe_kernel_bytes": (I, [P, C.c_char_p, C.POINTER(C.c_double)]),
    "sphe_set_tuning": (I, [P, I, I]),
    "sphe_get_tuning": (I, [P, C.POINTER(I), C.POINTER(I)]),
    "sphe_set_force_block": (I, [P, I]),
    "sphe_get_force_block": (I, [P, C.POINTER(I)]),
    "sphe_autotune": (I, [P, I]),
    "sphe_autotune_grid": (I, [P, I]),
    "sphe_set_scan_tile": (I, [P, I]),
    "sphe_set_persistent": (I, [P, I]),
    "sphe_get_persistent": (I, [P, C.POINTER(I)]),
    "sphe_set_sequence_steps": (I, [P, I]),
    "sphe_get_sequence_steps": (I, [P, C.POINTER(I)]),
    "sphe_get_scan_tile": (I, [P, C.POINTER(I)]),
    "sphe_set_lanes": (I, [P, I]),
    "sphe_get_lanes": (I, [P, C.POINTER(I)]),
    "sphe_save_state": (I, [P]),
    "sphe_restore_state": (I, [P]),
    "sphe_set_list_capacity": (I, [P, I]),
    "sphe_get_list_capacity": (I, [P, C.POINTER(I)]),
    "sphe_get_overflow_capacity": (I, [P, C.POINTER(I), C.POINTER(C.c_int64)]),
    "sphe_set_subcells": (I, [P, I]),
    "sphe_get_subcells": (I, [P, C.POINTER(I)]),
    "sphe_set_yz_split": (I, [P, I]),
    "sphe_get_yz_split": (I, [P, C.POINTER(I)]),
    "sphe_debug_cell_keys": (I, [P, I, P]),
    "sphe_debug_consts": (I, [P, I, P, I]),
    "sphe_debug_neighbours": (I, [P, I, P, P, P]),
    "sphe_debug_density": (I, [P, I, P]),
    "sphe_debug_accel": (I, [P, I, P]),
    "sphe_set_terrain_volume_fx": (I, [P, I, P, C.c_size_t]),
    "sphe_debug_copy": (I, [P, C.c_size_t, I, C.POINTER(C.c_double)]),
    "sphe_debug_pool_origin": (I, [P, C.c_int64]),
}

_lib = None


def lib():
    """The loaded library; raises if it was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libsphe.so not built: {LIB_PATH} (run make -C sph-erosion_amd)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


class SpheError(RuntimeError):
    pass


def check(rc, ctx=None, what=""):
    if rc != SPHE_OK:
        msg = lib().sphe_last_error(ctx).decode() if ctx else ""
        raise SpheError(f"{what}: {ERRORS.get(rc, rc)} {msg}")
    return rc
