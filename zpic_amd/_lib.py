"""Loader of the HIP library libzpic_em2d.so (built by __graft_entry__.build()).

There is no CPU fallback: importing the em2d API without the compiled
library raises, so every caller runs the GPU kernels or fails loudly.
"""
from __future__ import annotations

import ctypes as C
import os

from . import abi as A

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libzpic_em2d.so")

_lib = None


def load() -> C.CDLL:
    """The em2d GPU library with argument types declared."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"libzpic_em2d.so is not built ({LIB_PATH}); run __graft_entry__.build()")
    lib = C.CDLL(LIB_PATH)
    A.bind_api(lib, A.make_abi(True))
    P = C.POINTER
    ext = {
        "zpic_set_device": (None, [C.c_int]),
        "zpic_set_tiling": (None, [C.c_int, C.c_int, C.c_int]),
        "zpic_sim_set_open_y": (None, [C.c_void_p]),
        "zpic_set_device_init": (None, [C.c_int]),
        "zpic_device_init": (C.c_int, []),
        "zpic_set_device_inject": (None, [C.c_int]),
        "zpic_device_init_ambiguous": (C.c_longlong, []),
        "spec_deposit_charge_noalloc": (C.c_int, [C.c_void_p, C.POINTER(C.c_float)]),
        "zpic_sim_set_tiling": (None, [C.c_void_p, C.c_int, C.c_int, C.c_int]),
        "zpic_set_push_waves": (None, [C.c_int]),
        "zpic_set_fixed_deposit": (None, [C.c_int]),
        "zpic_set_window_shift_fusion": (None, [C.c_int]),
        "zpic_set_staged_sort": (None, [C.c_int]),
        "zpic_staged_sorts": (C.c_longlong, []),
        "zpic_staged_stats": (None, [P(C.c_longlong)]),
        "zpic_synchronize": (None, []),
        "zpic_rand_clear": (None, []),
        "zpic_rand_thread_private": (None, [C.c_int]),
        "spec_sync_host": (None, [C.c_void_p]),
        "emf_sync_host": (None, [C.c_void_p]),
        "current_sync_host": (None, [C.c_void_p]),
        "sim_sync_host": (None, [C.c_void_p]),
        "spec_mark_host_dirty": (None, [C.c_void_p]),
        "emf_mark_host_dirty": (None, [C.c_void_p]),
        "spec_track_ids": (None, [C.c_void_p]),
        "spec_get_ids": (None, [C.c_void_p, P(C.c_int64)]),
        "spec_overflow_count": (C.c_int, [C.c_void_p]),
        "zpic_kernel_timing": (None, [C.c_int]),
        "zpic_kernel_time": (C.c_double, [C.c_int, P(C.c_longlong)]),
        "zpic_kernel_timing_reset": (None, []),
        "zpic_batch_new": (C.c_void_p, [P(C.c_void_p), C.c_int]),
        "zpic_batch_run": (None, [C.c_void_p, C.c_int]),
        "zpic_batch_delete": (None, [C.c_void_p]),
    }
    for name, (res, args) in ext.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib
