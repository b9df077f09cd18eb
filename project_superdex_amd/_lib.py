"""Loader of libsdxfem.so (the HIP implementation).

There is no fallback: if the library is missing or fails to load, every
entry point raises.  Build it with ``python -m project_superdex_amd.build``.
"""
from __future__ import annotations

import ctypes as C
import os

from . import _abi

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsdxfem.so")

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int32)
_i64p = C.POINTER(C.c_int64)
_vp = C.c_void_p

# name -> (restype, argtypes); every symbol declared in include/sdx_fem.h
SIGNATURES = {
    "sdx_default_solver_params": (None, [C.POINTER(_abi.SolverParams)]),
    "sdx_default_soft_params": (None, [C.POINTER(_abi.SoftParams)]),
    "sdx_default_krylov": (None, [C.POINTER(_abi.KrylovParams)]),
    "sdx_last_error": (C.c_char_p, []),
    "sdx_abi_struct_size": (C.c_int32, [C.c_int32]),
    "sdx_pcg_bsr3_solve": (C.c_int32, [C.c_int32, C.c_int32, C.c_int32, _ip, _ip, _dp, C.c_int32, _dp, _dp,
                                       C.POINTER(_abi.KrylovParams), C.POINTER(_abi.LinearStatus)]),
    "sdx_pcg_bsr3_solve_f32": (C.c_int32, [C.c_int32, C.c_int32, C.c_int32, _ip, _ip, _fp, C.c_int32, _fp, _fp,
                                           C.POINTER(_abi.KrylovParams), C.POINTER(_abi.LinearStatus)]),
    "sdx_pcg_bsr3_solve_noalloc": (C.c_int32, [C.c_int32, C.c_int32, C.c_int32, _ip, _ip, _dp, C.c_int32, _dp,
                                               _dp, C.POINTER(_abi.KrylovParams), C.POINTER(_abi.LinearStatus)]),
    "sdx_world_create": (C.c_int32, [C.c_int32, C.c_int32, C.POINTER(_vp)]),
    "sdx_world_destroy": (C.c_int32, [_vp]),
    "sdx_add_soft_tet_mesh": (C.c_int32, [_vp, _dp, C.c_int32, _ip, C.c_int32, C.POINTER(_abi.SoftParams), _ip]),
    "sdx_set_element_materials": (C.c_int32, [_vp, C.c_int32, _dp, _dp, C.c_int32]),
    "sdx_pin_nodes": (C.c_int32, [_vp, C.c_int32, _ip, C.c_int32]),
    "sdx_add_rigid_plane": (C.c_int32, [_vp, _dp, _dp, C.c_int32, _ip]),
    "sdx_add_rigid_sphere": (C.c_int32, [_vp, _dp, C.c_double, C.c_int32, _ip]),
    "sdx_set_rigid_position": (C.c_int32, [_vp, C.c_int32, _dp]),
    "sdx_get_rigid_position": (C.c_int32, [_vp, C.c_int32, _dp]),
    "sdx_get_rigid_positions": (C.c_int32, [_vp, C.c_int32, _ip, _dp]),
    "sdx_set_rigid_positions": (C.c_int32, [_vp, C.c_int32, _ip, _dp]),
    "sdx_step": (C.c_int32, [_vp, C.c_double, C.POINTER(_abi.SolverParams), C.POINTER(_abi.StepStats), C.c_int32]),
    "sdx_get_displacements": (C.c_int32, [_vp, C.c_int32, _dp, C.c_int32]),
    "sdx_set_displacements": (C.c_int32, [_vp, C.c_int32, _dp, C.c_int32]),
    "sdx_get_velocities": (C.c_int32, [_vp, C.c_int32, _dp, C.c_int32]),
    "sdx_set_velocities": (C.c_int32, [_vp, C.c_int32, _dp, C.c_int32]),
    "sdx_get_positions_world": (C.c_int32, [_vp, C.c_int32, _dp, C.c_int32]),
    "sdx_capture_state": (C.c_int32, [_vp, _ip]),
    "sdx_restore_state": (C.c_int32, [_vp, C.c_int32, C.c_int32]),
    "sdx_release_state": (C.c_int32, [_vp, C.c_int32]),
    "sdx_is_equal_state": (C.c_int32, [_vp, C.c_int32, C.c_int32, _ip]),
    "sdx_state_size": (C.c_int32, [_vp, _i64p]),
    "sdx_capture_state_to_bytes": (C.c_int32, [_vp, C.c_void_p, C.c_int64]),
    "sdx_restore_state_from_bytes": (C.c_int32, [_vp, C.c_void_p, C.c_int64]),
    "sdx_capture_state_to_file": (C.c_int32, [_vp, C.c_char_p]),
    "sdx_restore_partial_state_from_bytes": (C.c_int32, [_vp, C.c_void_p, C.c_int64, C.c_uint32]),
    "sdx_is_equal_state_bytes": (C.c_int32, [C.c_void_p, C.c_int64, C.c_void_p, C.c_int64, _ip]),
    "sdx_state_bytes_to_json": (C.c_int32, [C.c_void_p, C.c_int64, C.c_int32, C.c_char_p, C.c_int64, _i64p]),
    "sdx_capture_state_to_json": (C.c_int32, [_vp, C.c_int32, C.c_char_p, C.c_int64, _i64p]),
    "sdx_world_info": (C.c_int32, [_vp, _i64p]),
    "sdx_actor_bsr_size": (C.c_int32, [_vp, C.c_int32, _ip, _ip]),
    "sdx_assemble": (C.c_int32, [_vp, C.c_int32, C.c_double, C.POINTER(_abi.SolverParams), _dp, _ip, _ip, _dp]),
    "sdx_objective": (C.c_int32, [_vp, C.c_int32, C.c_double, C.POINTER(_abi.SolverParams), _dp]),
    "sdx_set_boundary_conditions": (C.c_int32, [_vp, C.c_int32, _ip, _dp, C.c_int32, C.c_int32]),
    "sdx_register_contact_query": (C.c_int32, [_vp, C.c_int32]),
    "sdx_get_contact_wrench": (C.c_int32, [_vp, C.c_int32, C.c_int32, _dp]),
    "sdx_kernel_timing": (C.c_int32, [_vp, C.c_int32, C.c_int32, _dp]),
    "sdx_set_cg_burst": (C.c_int32, [_vp, C.c_int32]),
    "sdx_set_schedule": (C.c_int32, [_vp, C.POINTER(_abi.Schedule)]),
    "sdx_get_schedule": (C.c_int32, [_vp, C.POINTER(_abi.Schedule)]),
    "sdx_autotune": (C.c_int32, [_vp, C.c_double, C.POINTER(_abi.SolverParams), C.c_int32,
                                 C.POINTER(_abi.Schedule), _dp]),
    "sdx_default_train_params": (None, [C.POINTER(_abi.TrainParams)]),
    "sdx_mlp_create": (C.c_int32, [C.c_int32, C.c_uint64, C.POINTER(_vp)]),
    "sdx_mlp_destroy": (C.c_int32, [_vp]),
    "sdx_mlp_set_params": (C.c_int32, [_vp, _fp, C.c_int32]),
    "sdx_mlp_get_params": (C.c_int32, [_vp, _fp, C.c_int32]),
    "sdx_mlp_forward": (C.c_int32, [_vp, _fp, C.c_int64, C.c_int32, _fp]),
    "sdx_mlp_loss_grad": (C.c_int32, [_vp, _fp, _fp, C.c_int32, _fp, _dp]),
    "sdx_mlp_train": (C.c_int32, [_vp, _fp, _fp, C.c_int64, C.POINTER(_abi.TrainParams), _dp]),
    "sdx_mlp_train_collected": (C.c_int32, [_vp, _vp, C.POINTER(_abi.TrainParams), _dp]),
    "sdx_world_set_warm_start_model": (C.c_int32, [_vp, _vp]),
    "sdx_world_collect": (C.c_int32, [_vp, C.c_int64, C.c_int32]),
    "sdx_world_collected": (C.c_int32, [_vp, C.POINTER(C.c_int64), _fp, _fp, C.c_int64]),
}

_lib = None


class SdxError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"sdx error {code}: {msg}")
        self.code = code


def lib():
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m project_superdex_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        for idx, cls in enumerate([_abi.KrylovParams, _abi.LinearStatus, _abi.SolverParams, _abi.StepStats,
                                   _abi.SoftParams, _abi.Schedule, _abi.TrainParams]):
            if L.sdx_abi_struct_size(idx) != C.sizeof(cls):
                raise RuntimeError(f"ABI mismatch for {cls.__name__}: C {L.sdx_abi_struct_size(idx)} vs "
                                   f"Python {C.sizeof(cls)}")
        _lib = L
    return _lib


def check(code: int):
    if code != 0:
        raise SdxError(code, lib().sdx_last_error().decode())
    return code
