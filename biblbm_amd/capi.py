"""ctypes binding of libbiblbm.so (declarations in include/biblbm.h).

The library is required: importing this module without the built library
raises, there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbiblbm.so")

BL_OK = 0
STORE = {"f64": 0, "f32": 1, "f16s": 2}
KERNEL = {"roma": 0, "peskin": 1, "tophat": 2}
MODEL = {"capsule_skalak": 0, "capsule_neohookean": 1, "rbc": 2, "skalak_rbc": 3}

_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
_PI = ctypes.POINTER(ctypes.c_int)
_PLL = ctypes.POINTER(ctypes.c_longlong)
_PD = ctypes.POINTER(ctypes.c_double)

# name -> argtypes (all return int status, except bl_last_error)
SIGNATURES = {
    "bl_init": [_I],
    "bl_shutdown": [],
    "bl_device_info": [ctypes.c_char_p, _I, _PI, _PD],
    "bl_lattice_create": [_I, _I, _I, _P, _I, ctypes.POINTER(_P)],
    "bl_lattice_destroy": [_P],
    "bl_lattice_set_dynamics": [_P, _D, _D],
    "bl_lattice_set_mask": [_P, _P],
    "bl_lattice_set_skip_inert": [_P, _I],
    "bl_lattice_inert_cells": [_P, _PLL],
    "bl_lattice_set_velocity_walls": [_P, _I, _P, _P],
    "bl_lattice_set_body_force": [_P, _P],
    "bl_lattice_set_force_field": [_P, _P],
    "bl_lattice_set_phase": [_P, _P],
    "bl_lattice_init_equilibrium": [_P, _P, _P],
    "bl_lattice_put_populations": [_P, _P],
    "bl_lattice_get_populations": [_P, _P],
    "bl_lattice_initialize": [_P],
    "bl_collide_stream": [_P, _I],
    "bl_lattice_get_moments": [_P, _P, _P],
    "bl_lattice_get_force": [_P, _P],
    "bl_lattice_conserved": [_P, _P],
    "bl_lattice_conserved_static": [_P, _P],
    "bl_lattice_steps": [_P, _PLL],
    "bl_lattice_set_launch": [_P, _I, _I],
    "bl_lattice_set_launch_parity": [_P, _I, _I, _I, _I],
    "bl_lattice_get_launch": [_P, _I, _P],
    "bl_lattice_stream": [_P, ctypes.POINTER(_P)],
    "bl_lattice_time_collide": [_P, _I, _PD],
    "bl_sync": [],
    "bl_device_copy": [ctypes.c_longlong, _I, _PD],
    "bl_ibm_create": [_P, _I, ctypes.POINTER(_P)],
    "bl_ibm_destroy": [_P],
    "bl_shape_add": [_P, _I, _I, _P, _P, _PI],
    "bl_shape_info": [_P, _I, _P],
    "bl_particle_add": [_P, _I, _I, _P, _I, _P, _I, _PI],
    "bl_ibm_init": [_P],
    "bl_ibm_interpolate": [_P],
    "bl_ibm_compute_forces": [_P],
    "bl_ibm_spread": [_P],
    "bl_ibm_move": [_P],
    "bl_ibm_step": [_P],
    "bl_coupled_steps": [_P, _I, _I],
    "bl_coupled_prepare": [_P, _I],
    "bl_ibm_count": [_P, _PI, _PI],
    "bl_particle_get": [_P, _I, _I, _P],
    "bl_particle_set_positions": [_P, _I, _P],
    "bl_particle_set_state": [_P, _I, _I, _P],
    "bl_particle_set_params": [_P, _I, _P, _I],
    "bl_rbc_relax_nodes": [_P, _I, _D],
    "bl_particle_membrane_forces": [_P],
    "bl_ibm_set_vertex_forces": [_P, _P],
    "bl_particle_observables": [_P, _I, _P],
    "bl_ibm_get_phase": [_P, _P],
    "bl_ibm_stage_times": [_P, _P],
    "bl_ibm_tile_stats": [_P, _P],
    "bl_ibm_set_fused_staging": [_P, _I],
    "bl_ibm_set_launch": [_P, _P],
    "bl_ensemble_set_batched": [_P, _I],
    "bl_ensemble_is_batched": [_P, _PI],
    "bl_ensemble_tune": [_P, _I, ctypes.c_char_p, _I],
    "bl_ensemble_set_launch": [_P, _I, _I, _I],
    "bl_ibm_get_launch": [_P, _P],
    "bl_tune_ibm": [_P, _I, ctypes.c_char_p, _I],
    "bl_ibm_set_layered": [_P, _I],
    "bl_shape_set_tag": [_P, _I, ctypes.c_char_p],
    "bl_rbc_shrink_volume": [_P, _I, _D, _I, _D, _P],
    "bl_positions_create": [_P, _P, _D, ctypes.POINTER(_P)],
    "bl_positions_generate": [_P, ctypes.c_longlong, _D, ctypes.c_char],
    "bl_positions_get": [_P, ctypes.c_char, _P, ctypes.c_longlong, _PLL],
    "bl_positions_destroy": [_P],
    "bl_particle_pack": [_P, _I, _P, ctypes.c_longlong, _PLL],
    "bl_particle_describe": [_P, _I, _PI, _PI, _PLL],
    "bl_ibm_save_checkpoint": [_P, ctypes.c_char_p, ctypes.c_char_p],
    "bl_ibm_load_checkpoint": [_P, ctypes.c_char_p, ctypes.c_char_p],
    "bl_particle_write_vtk": [_P, _I, ctypes.c_char_p],
    "bl_particle_write_gmsh": [_P, _I, ctypes.c_char_p],
    "bl_ibm_write_lightweight": [_P, ctypes.c_char_p, ctypes.c_longlong],
    "bl_lattice_save_checkpoint": [_P, ctypes.c_char_p],
    "bl_lattice_load_checkpoint": [_P, ctypes.c_char_p],
    "bl_ensemble_create": [_I, _P, _P, _I, ctypes.POINTER(_P)],
    "bl_ensemble_destroy": [_P],
    "bl_ensemble_steps": [_P, _I, _I],
    "bl_ensemble_time_steps": [_P, _I, _I, _PD],
    "bl_tune_collide": [_P, ctypes.c_char_p, _I],
    "bl_tune_coupled": [_P, _I, ctypes.c_char_p, _I],
    "bl_tune_apply_coupled": [_P],
    "bl_tune_cache_save": [ctypes.c_char_p],
    "bl_tune_cache_load": [ctypes.c_char_p],
    "bl_guo_wall_create": [_P, _P, ctypes.POINTER(_P)],
    "bl_guo_wall_destroy": [_P],
    "bl_guo_wall_count": [_P, _PLL],
    "bl_guo_wall_nodes": [_P, _P, _P, _P, _P],
    "bl_guo_wall_set_velocity": [_P, _P],
    "bl_guo_wall_set_node_velocities": [_P, _P],
    "bl_guo_wall_insert": [_P],
    "bl_boundary_outside_mask": [_P, _I, _I, _I, _D, _P, ctypes.c_ubyte],
    "bl_ibm_set_boundary": [_P, _P],
    "bl_potential_eval": [_P, _D, _PD, _PD],
    "bl_ibm_add_wall_interaction": [_P, _P, _P],
    "bl_ibm_add_pp_force": [_P, _P],
    "bl_ibm_compute_collision_forces": [_P, _P, _P],
    "bl_rigid_add": [_P, _I, _P, _P],
    "bl_rigid_move": [_P],
    "bl_rigid_set_scale": [_P, _I, _D],
    "bl_rigid_get": [_P, _I, _P],
    "bl_rigid_grow": [_P, _I, _D, _I, _P, _P],
}


class BlError(RuntimeError):
    pass


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            "biblbm_amd: %s is not built; run `python __graft_entry__.py build` "
            "(there is no CPU fallback)" % LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH)
    lib.bl_last_error.restype = ctypes.c_char_p
    lib.bl_last_error.argtypes = []
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name, None)
        if fn is None:
            continue
        fn.restype = ctypes.c_int
        fn.argtypes = args
    return lib


lib = _load()


def exported_symbols():
    return [n for n in list(SIGNATURES) + ["bl_last_error"] if hasattr(lib, n)]


def check(rc, what=""):
    if rc != BL_OK:
        msg = lib.bl_last_error().decode(errors="replace")
        raise BlError("%s failed (status %d): %s" % (what, rc, msg))


def call(name, *args):
    check(getattr(lib, name)(*args), name)


def ptr(a):
    return None if a is None else a.ctypes.data_as(_P)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def u8(a):
    return np.ascontiguousarray(a, dtype=np.uint8)
