"""ctypes binding of the C ABI in include/swpc_hip.h (libswpc_hip.so).

The library is built in-tree by ``openswpc_amd.build`` (or
``__graft_entry__.build()``).  Loading fails loudly when it is missing:
there is no CPU fallback on the product path.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libswpc_hip.so"

SWPC_FP32, SWPC_FP64, SWPC_FP32_S16, SWPC_FP32_R16 = 0, 1, 2, 3

P = C.c_void_p
I = C.c_int
D = C.c_double
PI_ = C.POINTER(C.c_int)
PF = C.POINTER(C.c_float)
PD = C.POINTER(C.c_double)
PLL = C.POINTER(C.c_longlong)


class Desc3(C.Structure):
    _fields_ = [("nx", I), ("ny", I), ("nz", I), ("na", I), ("nm", I), ("precision", I),
                ("dx", D), ("dy", D), ("dz", D), ("dt", D), ("ts", C.c_float * 8)]


class Desc2(C.Structure):
    _fields_ = [("nx", I), ("nz", I), ("na", I), ("nm", I), ("precision", I),
                ("dx", D), ("dz", D), ("dt", D), ("ts", C.c_float * 8)]


# name -> (restype, argtypes); every entry point returns int status
_PROTOS = {
    "swpc_last_error": (C.c_char_p, []),
    "swpc_version": (I, []),
    "swpc_device_count": (I, []),
    "swpc_membw": (I, [C.c_longlong, I, PD]),
    "swpc_rmedia_generate": (I, [I, PI_, PD, PD, D, D, I, PF, PF]),
    "swpc3d_create": (I, [C.POINTER(Desc3), C.POINTER(P)]),
    "swpc3d_destroy": (I, [P]),
    "swpc3d_set_medium": (I, [P] + [PF] * 10),
    "swpc3d_set_density": (I, [P, PF]),
    "swpc3d_setup_medium": (I, [P, PF, PF, PF, PF, PF, C.c_float, C.c_float, C.c_float, PI_, PI_, PF]),
    "swpc3d_get_medium": (I, [P, I, PF]),
    "swpc3d_set_surface": (I, [P] + [PI_] * 5),
    "swpc3d_set_pml": (I, [P] + [PF] * 6),
    "swpc3d_set_sources": (I, [P, I, PI_, PI_, PI_, PD, I, PD]),
    "swpc3d_set_bodyforce": (I, [P, I, PI_, PI_, PI_, PD, I, PD]),
    "swpc3d_set_receivers": (I, [P, I, PI_, PI_, PI_, I, I]),
    "swpc3d_update_stress": (I, [P]),
    "swpc3d_update_vel": (I, [P]),
    "swpc3d_stressglut": (I, [P, I]),
    "swpc3d_bodyforce": (I, [P, I]),
    "swpc3d_store_receivers": (I, [P, I]),
    "swpc3d_add_snapshot": (I, [P, I, I, PI_, I, I, C.c_float, C.c_float, C.c_float, PI_]),
    "swpc3d_snapshots": (I, [P, I]),
    "swpc3d_get_snapshot": (I, [P, I, I, PF]),
    "swpc3d_snapshot_peaks": (I, [P, I]),
    "swpc3d_finish_snapshots": (I, [P, I]),
    "swpc3d_get_peaks": (I, [P, I, PF]),
    "swpc3d_step": (I, [P, I]),
    "swpc3d_run": (I, [P, I, I]),
    "swpc3d_get_receivers": (I, [P, PF, PD]),
    "swpc3d_get_field": (I, [P, I, PD]),
    "swpc3d_set_field": (I, [P, I, PD]),
    "swpc3d_vmax": (I, [P, PF]),
    "swpc3d_sync": (I, [P]),
    "swpc3d_profile": (I, [P, I]),
    "swpc3d_profile_read": (I, [P, PD, PLL]),
    "swpc3d_memory": (I, [P, PLL]),
    "swpc3d_set_tuning": (I, [P, PI_, I]),
    "swpc_lab_nparam": (I, []),
    "swpc3d_set_lab": (I, [P, PF]),
    "swpc3d_set_lab_sponge": (I, [P, C.c_float]),
    "swpc_lab_trainer_create": (I, [I, C.POINTER(P)]),
    "swpc_lab_trainer_destroy": (I, [P]),
    "swpc_lab_begin_run": (I, [P, P, I, I, I, I, I, PI_]),
    "swpc3d_lab_gather": (I, [P, P, I, C.c_float, C.c_uint, C.c_float]),
    "swpc_lab_count": (I, [P, PI_]),
    "swpc_lab_add": (I, [P, I, I, PF]),
    "swpc_lab_samples": (I, [P, I, I, I, PF]),
    "swpc_lab_train": (I, [P, I, PF, I, I, C.c_float, C.c_float, C.c_uint, PF]),
    "swpc_lab_grad": (I, [P, PF, I, I, PF, PD, PF]),
    "swpc_lab_grad_noalloc": (I, [PF, I, I, PF, PD, PF]),
    "swpc_batch_create": (I, [C.POINTER(P)]),
    "swpc_batch_destroy": (I, [P]),
    "swpc_batch_add3d": (I, [P, P]),
    "swpc_batch_add2d": (I, [P, P]),
    "swpc_batch_run": (I, [P, I, I, I]),
    "swpc_batch_sync": (I, [P]),
    "swpc_batch_set_grouping": (I, [P, I]),
    "swpcsh_create": (I, [C.POINTER(Desc2), C.POINTER(P)]),
    "swpcsh_destroy": (I, [P]),
    "swpcsh_set_medium": (I, [P, PF, PF, PF]),
    "swpcsh_set_surface": (I, [P, PI_, PI_, PI_, PI_, PI_]),
    "swpcsh_set_pml": (I, [P, PF, PF, PF, PF]),
    "swpcsh_set_sources": (I, [P, I, PI_, PI_, PD, I, PD]),
    "swpcsh_set_bodyforce": (I, [P, I, PI_, PI_, PD, I, PD]),
    "swpcsh_set_receivers": (I, [P, I, PI_, PI_, I, I]),
    "swpcsh_update_stress": (I, [P]),
    "swpcsh_update_vel": (I, [P]),
    "swpcsh_stressglut": (I, [P, I]),
    "swpcsh_bodyforce": (I, [P, I]),
    "swpcsh_store_receivers": (I, [P, I]),
    "swpcsh_add_snapshot": (I, [P, I, I, PI_, I, I, C.c_float, C.c_float, C.c_float, PI_]),
    "swpcsh_snapshots": (I, [P, I]),
    "swpcsh_get_snapshot": (I, [P, I, I, PF]),
    "swpcsh_step": (I, [P, I]),
    "swpcsh_set_graph": (I, [P, I]),
    "swpcsh_run": (I, [P, I, I]),
    "swpcsh_get_receivers": (I, [P, PF, PD]),
    "swpcsh_get_field": (I, [P, I, PD]),
    "swpcsh_set_field": (I, [P, I, PD]),
    "swpcsh_vmax": (I, [P, PF]),
    "swpcsh_sync": (I, [P]),
    "swpc2d_create": (I, [C.POINTER(Desc2), C.POINTER(P)]),
    "swpc2d_destroy": (I, [P]),
    "swpc2d_set_medium": (I, [P] + [PF] * 5),
    "swpc2d_set_surface": (I, [P] + [PI_] * 5),
    "swpc2d_set_pml": (I, [P] + [PF] * 4),
    "swpc2d_set_sources": (I, [P, I, PI_, PI_, PD, I, PD]),
    "swpc2d_set_bodyforce": (I, [P, I, PI_, PI_, PD, I, PD]),
    "swpc2d_set_receivers": (I, [P, I, PI_, PI_, I, I]),
    "swpc2d_update_stress": (I, [P]),
    "swpc2d_update_vel": (I, [P]),
    "swpc2d_stressglut": (I, [P, I]),
    "swpc2d_bodyforce": (I, [P, I]),
    "swpc2d_store_receivers": (I, [P, I]),
    "swpc2d_add_snapshot": (I, [P, I, I, PI_, I, I, C.c_float, C.c_float, C.c_float, PI_]),
    "swpc2d_snapshots": (I, [P, I]),
    "swpc2d_get_snapshot": (I, [P, I, I, PF]),
    "swpc2d_set_graph": (I, [P, I]),
    "swpc2d_set_fusion": (I, [P, I]),
    "swpc2d_step": (I, [P, I]),
    "swpc2d_run": (I, [P, I, I]),
    "swpc2d_get_receivers": (I, [P, PF, PD]),
    "swpc2d_get_field": (I, [P, I, PD]),
    "swpc2d_set_field": (I, [P, I, PD]),
    "swpc2d_vmax": (I, [P, PF]),
    "swpc2d_sync": (I, [P]),
}

_lib = None


class SwpcError(RuntimeError):
    pass


def lib():
    """Load libswpc_hip.so once; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise SwpcError(f"{LIB_PATH} is missing: run `python -m openswpc_amd.build` "
                            "(or __graft_entry__.build()) first; there is no CPU fallback")
        L = C.CDLL(str(LIB_PATH))
        for name, (res, args) in _PROTOS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc, what=""):
    if rc != 0:
        msg = lib().swpc_last_error().decode(errors="replace")
        raise SwpcError(f"{what} failed ({rc}): {msg}")
    return rc


def ptr(a, ctype):
    """Pointer to a contiguous numpy array (None passes NULL)."""
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(ctype))
