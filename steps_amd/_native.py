"""ctypes binding of the engine's C ABI (include/stg.h).

The engine library is built in-tree at steps_amd/_lib/libstg.so by
``python -c "import __graft_entry__ as g; g.build()"`` (or
``make -C steps_amd/csrc``). Loading fails loudly when the library is missing:
there is no CPU fallback for the solver.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "_lib" / "libstg.so"

STG_OK = 0
STG_ERR_ARG = 1
STG_ERR_NOTIMPL = 2
STG_ERR_PROG = 3
STG_ERR_SYS = 4
STG_ERR_DEVICE = 5

MODE_TETOPSPLIT = 0
MODE_DISTTETOPSPLIT = 1

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class StgMesh(C.Structure):
    _fields_ = [
        ("n_tets", C.c_int32),
        ("tet_vol", _dp),
        ("tet_nbr", _ip),
        ("tet_nbr_face", _ip),
        ("tet_face_g", _dp),
        ("n_tris", C.c_int32),
        ("tri_tet", _ip),
        ("tri_area", _dp),
        ("tet_nbr_comp", _ip),
        ("tet_face_bdir", _ip),
        ("tri_nbr", _ip),
        ("tri_nbr_edge", _ip),
        ("tri_edge_g", _dp),
        ("tri_otet", _ip),
        ("tri_ocomp", _ip),
        ("tri_nbr_comp", _ip),
    ]


class StgModel(C.Structure):
    _fields_ = [
        ("n_specs", C.c_int32),
        ("n_reacs", C.c_int32),
        ("reac_lhs", _ip),
        ("reac_upd", _ip),
        ("reac_k", _dp),
        ("spec_dcst", _dp),
        ("n_sspecs", C.c_int32),
        ("n_sreacs", C.c_int32),
        ("sreac_lhs_s", _ip),
        ("sreac_upd_s", _ip),
        ("sreac_lhs_i", _ip),
        ("sreac_upd_i", _ip),
        ("sreac_k", _dp),
        ("sreac_vdep", _ip),
        ("n_vdep", C.c_int32),
        ("vdep_vmin", _dp),
        ("vdep_dv", _dp),
        ("vdep_off", _ip),
        ("vdep_k", _dp),
        ("n_ghk", C.c_int32),
        ("ghk_chan", _ip),
        ("ghk_ion", _ip),
        ("ghk_valence", _ip),
        ("ghk_perm", _dp),
        ("ghk_voconc", _dp),
        ("ghk_vshift", _dp),
        ("ghk_realflux", _ip),
        ("sspec_dcst", _dp),
        ("sreac_lhs_o", _ip),
        ("sreac_upd_o", _ip),
    ]


class StgRealization(C.Structure):
    _fields_ = [
        ("seed", C.c_uint64),
        ("stream", C.c_uint32),
        ("reac_k", _dp),
        ("sreac_k", _dp),
        ("spec_dcst", _dp),
    ]


class StgEfieldMesh(C.Structure):
    _fields_ = [
        ("n_verts", C.c_int32),
        ("verts", _dp),
        ("n_tets", C.c_int32),
        ("tets", _ip),
        ("n_tris", C.c_int32),
        ("tris", _ip),
    ]


# name -> (restype, argtypes); the list is the header's entry points, in order.
SIGNATURES = {
    "stg_create": (C.c_int, [C.POINTER(StgModel), C.POINTER(StgMesh), C.c_uint64, C.c_int32,
                             C.POINTER(C.c_void_p)]),
    "stg_create_batch": (C.c_int, [C.POINTER(StgModel), C.c_int32, C.POINTER(StgMesh),
                                   C.POINTER(StgRealization), C.c_int32, C.POINTER(C.c_void_p)]),
    "stg_destroy": (C.c_int, [C.c_void_p]),
    "stg_last_error": (C.c_char_p, [C.c_void_p]),
    "stg_reset": (C.c_int, [C.c_void_p]),
    "stg_run": (C.c_int, [C.c_void_p, C.c_double]),
    "stg_advance": (C.c_int, [C.c_void_p, C.c_double]),
    "stg_set_seed": (C.c_int, [C.c_void_p, C.c_int32, C.c_uint64]),
    "stg_get_time": (C.c_int, [C.c_void_p, _dp]),
    "stg_set_time": (C.c_int, [C.c_void_p, C.c_double]),
    "stg_get_nsteps": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_uint64)]),
    "stg_set_nsteps": (C.c_int, [C.c_void_p, C.c_int32, C.c_uint64]),
    "stg_get_update_period": (C.c_int, [C.c_void_p, C.c_int32, _dp]),
    "stg_get_extents": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_uint64),
                                  C.POINTER(C.c_uint64)]),
    "stg_set_tet_counts": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, _ip]),
    "stg_get_tet_counts": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, _ip]),
    "stg_set_tri_counts": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, _ip]),
    "stg_get_tri_counts": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, _ip]),
    "stg_get_batch_tet_counts": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32,
                                           C.POINTER(C.c_int64), C.c_size_t, _dp]),
    "stg_sum_tet_counts": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.POINTER(C.c_int64)]),
    "stg_sum_tri_counts": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.POINTER(C.c_int64)]),
    "stg_sum_all_tet_counts": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64)]),
    "stg_sum_all_tri_counts": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64)]),
    "stg_set_reac_k": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_double]),
    "stg_set_sreac_k": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_double]),
    "stg_set_dcst": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_double]),
    "stg_set_launch": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32]),
    "stg_set_lanes": (C.c_int, [C.c_void_p, C.c_int32]),
    "stg_get_lanes": (C.c_int, [C.c_void_p, _ip]),
    "stg_set_slot_width": (C.c_int, [C.c_void_p, C.c_int32]),
    "stg_get_slot_width": (C.c_int, [C.c_void_p, _ip]),
    "stg_set_specialize": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32]),
    "stg_get_specialize": (C.c_int, [C.c_void_p, _ip]),
    "stg_set_diff_apply_threshold": (C.c_int, [C.c_void_p, C.c_int32]),
    "stg_state_size": (C.c_int, [C.c_void_p, C.POINTER(C.c_size_t)]),
    "stg_checkpoint": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "stg_restore": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "stg_set_tri_v": (C.c_int, [C.c_void_p, C.c_int32, _dp]),
    "stg_get_tri_v": (C.c_int, [C.c_void_p, C.c_int32, _dp]),
    "stg_set_temp": (C.c_int, [C.c_void_p, C.c_double]),
    "stg_create_comps": (C.c_int, [C.POINTER(StgModel), C.c_int32, C.POINTER(StgMesh),
                                   C.POINTER(StgRealization), C.c_int32, C.POINTER(C.c_void_p)]),
    "stg_sum_batch_tet_counts": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.POINTER(C.c_int64),
                                           C.c_size_t, C.POINTER(C.c_int64)]),
    "stg_get_batch_tet_counts_noalloc": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32,
                                                   C.POINTER(C.c_int64), C.c_size_t, _dp]),
    "stg_set_sdcst": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_double]),
    "stg_set_boundary_dcst": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_double]),
    "stg_get_boundary_dcst": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, _dp]),
    "stg_get_temp": (C.c_int, [C.c_void_p, _dp]),
    "stg_get_sizes": (C.c_int, [C.c_void_p, C.c_int32, _ip, _ip, _ip]),
    "stg_last_run_kernel_ms": (C.c_int, [C.c_void_p, _dp, _ip]),
    "stg_set_tet_reac_table": (C.c_int, [C.c_void_p, C.c_int32, _dp]),
    "stg_set_tet_diff_table": (C.c_int, [C.c_void_p, C.c_int32, _dp]),
    "stg_set_tet_clamp": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_uint32)]),
    "stg_set_tri_sreac_table": (C.c_int, [C.c_void_p, C.c_int32, _dp]),
    "stg_set_tri_sdiff_table": (C.c_int, [C.c_void_p, C.c_int32, _dp]),
    "stg_set_tri_clamp": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_uint32)]),
    "stg_efield_create": (C.c_int, [C.POINTER(StgEfieldMesh), C.POINTER(C.c_void_p)]),
    "stg_efield_destroy": (C.c_int, [C.c_void_p]),
    "stg_efield_last_error": (C.c_char_p, [C.c_void_p]),
    "stg_efield_set_potential": (C.c_int, [C.c_void_p, C.c_double]),
    "stg_efield_set_capac": (C.c_int, [C.c_void_p, C.c_double]),
    "stg_efield_set_volres": (C.c_int, [C.c_void_p, C.c_double]),
    "stg_efield_set_memb_res": (C.c_int, [C.c_void_p, C.c_double, C.c_double]),
    "stg_efield_set_vert_iclamp": (C.c_int, [C.c_void_p, C.c_int32, C.c_double]),
    "stg_efield_set_tri_iclamp": (C.c_int, [C.c_void_p, C.c_int32, C.c_double]),
    "stg_efield_set_vert_clamped": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32]),
    "stg_efield_set_verts_v": (C.c_int, [C.c_void_p, _dp]),
    "stg_efield_get_verts_v": (C.c_int, [C.c_void_p, _dp]),
    "stg_efield_get_tris_v": (C.c_int, [C.c_void_p, _dp]),
    "stg_efield_set_ohmic": (C.c_int, [C.c_void_p, C.c_int32, _ip, _dp, _dp]),
    "stg_efield_set_tri_ohmic_erev": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_double]),
    "stg_efield_get_tri_ohmic_erev": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, _dp]),
    "stg_efield_set_tri_capac": (C.c_int, [C.c_void_p, C.c_int32, C.c_double]),
    "stg_efield_get_tri_capac": (C.c_int, [C.c_void_p, C.c_int32, _dp]),
    "stg_efield_get_tri_ghk_currents": (C.c_int, [C.c_void_p, C.c_int32, _dp]),
    "stg_efield_attach": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32]),
    "stg_efield_advance": (C.c_int, [C.c_void_p, C.c_double]),
    "stg_efield_run": (C.c_int, [C.c_void_p, C.c_void_p, C.c_double, C.c_double]),
    "stg_efield_set_solver": (C.c_int, [C.c_void_p, C.c_int32, C.c_double, C.c_int32]),
    "stg_efield_stats": (C.c_int, [C.c_void_p, _ip, _ip, _dp]),
    "stg_efield_set_precond": (C.c_int, [C.c_void_p, C.c_int32]),
    "stg_efield_precond_info": (C.c_int, [C.c_void_p, _ip, _ip, _dp]),
    "stg_efield_sizes": (C.c_int, [C.c_void_p, _ip, _ip, _ip]),
    "stg_efield_state_size": (C.c_int, [C.c_void_p, C.POINTER(C.c_size_t)]),
    "stg_efield_checkpoint": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "stg_efield_restore": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "stg_mesh_build": (C.c_int, [_dp, C.c_int32, _ip, C.c_int32, _ip, C.c_int32, C.POINTER(C.c_void_p)]),
    "stg_mesh_build_sizes": (C.c_int, [C.c_void_p, _ip, _ip]),
    "stg_mesh_build_fetch": (C.c_int, [C.c_void_p, _dp, _dp, _ip, _ip, _ip, _dp, _dp, _ip, _ip, _dp, _dp]),
    "stg_mesh_build_destroy": (C.c_int, [C.c_void_p]),
    "stg_mesh_build_last_error": (C.c_char_p, []),
}


class StgError(RuntimeError):
    """Engine failure; ``code`` is the C status, ``kind`` the matching STEPS error."""

    KINDS = {STG_ERR_ARG: "ArgErr", STG_ERR_NOTIMPL: "NotImplErr", STG_ERR_PROG: "ProgErr",
             STG_ERR_SYS: "SysErr", STG_ERR_DEVICE: "DeviceErr"}

    def __init__(self, code: int, msg: str):
        self.code = code
        self.kind = self.KINDS.get(code, "Err")
        super().__init__(f"{self.kind}: {msg}")


class ArgErr(StgError, ValueError):
    pass


class NotImplErr(StgError, NotImplementedError):
    pass


_ERR_CLASS = {STG_ERR_ARG: ArgErr, STG_ERR_NOTIMPL: NotImplErr}

_lib = None


def load(path: os.PathLike | None = None) -> C.CDLL:
    """Load the engine library (once). Raises if it has not been built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path is not None else LIB_PATH
    if not p.exists():
        raise ImportError(
            f"steps_amd engine library not found at {p}; build it with "
            "`python -c 'import __graft_entry__ as g; g.build()'` (hipcc, gfx950)")
    lib = C.CDLL(str(p))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def check(rc: int, handle=None, efield: bool = False) -> None:
    if rc == STG_OK:
        return
    lib = load()
    msg = (lib.stg_efield_last_error if efield else lib.stg_last_error)(handle)
    msg = msg.decode() if msg else "unknown error"
    raise _ERR_CLASS.get(rc, StgError)(rc, msg)


def ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


def as_i32(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.int32))


def as_f64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))
