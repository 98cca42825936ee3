"""Loader of the native libraries built in-tree under sayakaamr_amd/lib/.

libsayaka_host.so  host tree, tables and the solver driver (C ABI: include/sayaka_host.h)
libsayaka_dev.so   HIP kernels for gfx950 and the device C ABI (include/sayaka_amd.h)

There is no fallback: a missing library raises, so a solve can never silently run
anywhere but on the GPU.
"""
from __future__ import annotations

import atexit
import ctypes as C
import os
import sys
import weakref

HERE = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.path.join(HERE, "lib")
HOST_LIB = os.path.join(LIBDIR, "libsayaka_host.so")
DEV_LIB = os.path.join(LIBDIR, "libsayaka_dev.so")
DEV_PHASES_LIB = os.path.join(LIBDIR, "libsayaka_dev_phases.so")   # the phase-timing build

i32, f64, vp = C.c_int, C.c_double, C.c_void_p
_host = None


class NativeLibraryMissing(RuntimeError):
    pass


def ptr(a):
    return None if a is None else a.ctypes.data_as(vp)


_HOST_SIG = {
    "sy_tree_create": ([i32, i32, i32, i32, i32, i32, vp, vp, vp, vp], vp),
    "sy_tree_destroy": ([vp], None),
    "sy_tree_numblocks": ([vp], i32),
    "sy_tree_finest": ([vp], i32),
    "sy_tree_set_flags": ([vp, vp, vp, i32], i32),
    "sy_tree_get_flags": ([vp, vp, vp, i32], i32),
    "sy_tree_regrid": ([vp], i32),
    "sy_tree_regrid_origin": ([vp, vp], None),
    "sy_tree_flag_shell": ([vp, vp, f64, f64], i32),
    "sy_tree_export": ([vp] + [vp] * 14, None),
    "sy_tree_level_count": ([vp, i32], i32),
    "sy_tree_level_blocks": ([vp, i32, vp], None),
    "sy_tree_build_shell": ([vp, vp, f64, f64], i32),
    "sy_mac_last_error": ([], C.c_char_p),
    "sy_mac_create": ([vp, i32], vp),
    "sy_mac_destroy": ([vp], None),
    "sy_mac_prepare": ([vp], i32),
    "sy_mac_set_bc": ([vp, i32, vp, vp], i32),
    "sy_mac_set_ib_ls": ([vp, vp, i32], i32),
    "sy_mac_set_ib_frac": ([vp, vp, vp], i32),
    "sy_mac_set_coefb": ([vp, vp], i32),
    "sy_mac_set_rhs": ([vp, vp], i32),
    "sy_mac_set_init_sol": ([vp, vp], i32),
    "sy_mac_fix_bc": ([vp, vp, i32, vp], i32),
    "sy_mac_set_zero_sol": ([vp], i32),
    "sy_mac_get_sol": ([vp, vp], i32),
    "sy_mac_flux": ([vp, i32, i32, vp, vp], i32),
    "sy_mac_set_params": ([vp, f64, f64, i32, i32, i32, i32, i32], i32),
    "sy_mac_cg_solve": ([vp, vp, vp, vp, vp], i32),
    "sy_mac_mg_solve": ([vp, i32, vp, vp, vp, vp], i32),
    "sy_mac_relax_solve": ([vp, vp, vp, vp, vp], i32),
    "sy_writer_last_error": ([], C.c_char_p),
    "sy_tree_last_error": ([], C.c_char_p),
    "sy_tree_state_register": ([vp, i32, i32, i32, vp, i32, vp, vp, vp], i32),
    "sy_tree_state_set_grown": ([vp, i32, vp], i32),
    "sy_tree_state_get_grown": ([vp, i32, vp], i32),
    "sy_tree_init_new_data": ([vp], i32),
    "sy_tree_state_fill": ([vp, i32, i32, i32, i32], i32),
    "sy_tree_state_set_resident": ([vp, i32, i32, i32], i32),
    "sy_tree_transfer_stats": ([vp, vp], i32),
    "sy_write_chombo": ([vp, i32, vp, C.c_char_p, f64, f64, i32, vp], i32),
}


def host():
    """The host library (tree, tables, solver driver); raises if it was not built."""
    global _host
    if _host is not None:
        return _host
    if not os.path.exists(HOST_LIB):
        raise NativeLibraryMissing(f"{HOST_LIB} not built; run __graft_entry__.build()")
    lib = C.CDLL(HOST_LIB)
    for name, (args, res) in _HOST_SIG.items():
        fn = getattr(lib, name, None)
        if fn is None:
            continue
        fn.argtypes = args
        fn.restype = res
    _host = lib
    return lib


def register(lib, sig: dict):
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res


# ------------------------------------------------------------------ orderly teardown
# Objects holding a device context or device memory register here.  At interpreter exit the
# atexit hook closes the live ones, newest first, while the HIP runtime is still fully up; a
# finalizer that runs later (during interpreter finalization) makes no HIP call at all.
_live = {}
_next_id = 0


def track(obj):
    """Register obj (with a close() method) for closing at exit; returns its key."""
    global _next_id
    _next_id += 1
    _live[_next_id] = weakref.ref(obj)
    return _next_id


def untrack(key):
    _live.pop(key, None)


def finalizing():
    return sys.is_finalizing()


def _close_all():
    for key in sorted(_live, reverse=True):
        ref = _live.pop(key, None)
        obj = ref() if ref is not None else None
        if obj is not None:
            try:
                obj.close()
            except Exception:
                pass


atexit.register(_close_all)
