"""ctypes binding of libtess2_amd.so (the C ABI in include/tess2_amd.h).

The library is built in-tree by tess2_amd.build; importing this module never
falls back to anything else: a missing library raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtess2_amd.so")

TESS2_DENSE_TESS, TESS2_DENSE_CIC = 0, 1
ASSIGN_REFERENCE_SCAN, ASSIGN_EXACT_VORONOI = 0, 1
PREC_MIXED, PREC_FP64 = 0, 1

ERRORS = {0: "ok", -1: "invalid argument", -2: "HIP error", -3: "capacity exceeded",
          -4: "out of device memory", -5: "I/O error",
          -6: "internal consistency check failed"}


class Tess2Error(RuntimeError):
    pass


class Block(C.Structure):
    """tess2_block_t"""
    _fields_ = [("gid", C.c_int), ("num_orig_particles", C.c_int),
                ("num_particles", C.c_int), ("particles", C.c_void_p),
                ("num_tets", C.c_int), ("tets", C.c_void_p),
                ("vert_to_tet", C.c_void_p), ("bounds_min", C.c_float * 3),
                ("bounds_max", C.c_float * 3), ("ghost_gid", C.c_void_p),
                ("ghost_cell", C.c_void_p)]


class Stats(C.Structure):
    """tess2_stats_t"""
    _fields_ = [("cells_scan", C.c_longlong), ("cells_cic", C.c_longlong),
                ("cells_incomplete", C.c_longlong), ("cells_outside", C.c_longlong),
                ("cells_degenerate", C.c_longlong), ("points_deposited", C.c_longlong),
                ("points_offblock", C.c_longlong), ("points_routed", C.c_longlong),
                ("points_shared", C.c_longlong), ("fp64_fallbacks", C.c_longlong),
                ("tot_mass", C.c_double), ("max_dense", C.c_float),
                ("device_ms", C.c_float), ("cells_deferred", C.c_longlong)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None

F32P = C.POINTER(C.c_float)
I32P = C.POINTER(C.c_int)

# exported symbol -> (restype, argtypes)
SIGNATURES = {
    "tess2_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int]),
    "tess2_destroy": (C.c_int, [C.c_void_p]),
    "tess2_last_error": (C.c_char_p, [C.c_void_p]),
    "tess2_fill_circumcenters": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p,
                                           C.c_int, C.c_void_p]),
    "tess2_volumes": (C.c_int, [C.c_void_p, C.POINTER(Block), C.c_void_p]),
    "tess2_ghost_sends": (C.c_int, [C.c_void_p, C.POINTER(Block), C.c_int, C.c_void_p, C.c_int,
                                    C.c_void_p]),
    "tess2_volumes_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_void_p,
                                        C.c_int, C.c_void_p, C.c_void_p]),
    "tess2_grid_step_params": (None, [C.c_int, F32P, F32P, F32P, F32P, F32P, F32P, F32P,
                                      I32P]),
    "tess2_block_grid_params": (None, [F32P, F32P, I32P, I32P, I32P, F32P, F32P, C.c_float,
                                       F32P, F32P, I32P]),
    "tess2_dense": (C.c_int, [C.c_void_p, C.c_int, C.c_int, F32P, F32P, C.c_int, F32P,
                              C.c_float, F32P, F32P, F32P, F32P, F32P, C.c_float, I32P,
                              C.POINTER(Block), C.c_int, C.c_int, C.c_int,
                              C.POINTER(F32P), C.POINTER(Stats)]),
    "tess2_dense_cells": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]),
    "tess2_write_grid": (C.c_int, [C.c_char_p, C.c_void_p, I32P, C.c_int]),
    "tess2_project_grid": (C.c_int, [C.c_int, C.c_void_p, C.POINTER(F32P), C.c_void_p,
                                     C.c_void_p]),
    "tess2_write_blocks": (C.c_int, [C.c_char_p, C.c_int, C.c_void_p, C.POINTER(F32P),
                                     C.c_void_p, C.c_void_p, I32P, C.c_int]),
    "tess2_stage": (C.c_int, [C.c_void_p, C.POINTER(Block), C.c_int]),
    "tess2_run": (C.c_int, [C.c_void_p, C.c_int, C.c_int, F32P, F32P, C.c_int, F32P,
                            C.c_float, C.c_float, I32P, C.c_int, C.c_int, C.c_int]),
    "tess2_sync": (C.c_int, [C.c_void_p, C.POINTER(Stats)]),
    "tess2_fetch": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]),
    "tess2_kernel_time": (C.c_int, [C.c_void_p, F32P, I32P]),
    "tess2_fetch_owner": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p]),
    "tess2_set_option": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    "tess2_get_option": (C.c_int, [C.c_void_p, C.c_int, I32P]),
    "tess2_autotune": (C.c_int, [C.c_void_p, C.c_int, F32P]),
}

OPT_GRAPH, OPT_WAVES_PER_CU, OPT_CHUNK_SMALL, OPT_CHUNK_MEDIUM, OPT_CHUNK_LARGE = 0, 1, 2, 3, 4
OPT_CELL_KERNEL = 5


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise Tess2Error(f"{LIB_PATH} is not built; run tess2_amd.build.build()")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def f3(x):
    a = (C.c_float * 3)()
    if x is not None:
        a[:] = [float(np.float32(v)) for v in x]
    return a


def i3(x):
    a = (C.c_int * 3)()
    a[:] = [int(v) for v in x]
    return a


def fptr(a: np.ndarray):
    return a.ctypes.data_as(F32P)


def iptr(a: np.ndarray):
    return a.ctypes.data_as(I32P)


class Context:
    """One device context (library stream + staged blocks)."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        r = lib().tess2_create(C.byref(h), int(device))
        if r != 0:
            raise Tess2Error(f"tess2_create failed: {ERRORS.get(r, r)}")
        self.h = h
        self._keep = []

    def check(self, r, what):
        if r != 0:
            msg = lib().tess2_last_error(self.h)
            raise Tess2Error(f"{what}: {ERRORS.get(r, r)}: "
                             f"{msg.decode() if msg else ''}")

    def close(self):
        if self.h:
            lib().tess2_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def make_block(hb) -> tuple[Block, list]:
    """tess2_block_t view of a HostBlock (keeps the arrays alive)."""
    p = np.ascontiguousarray(hb.particles, dtype=np.float32)
    t = np.ascontiguousarray(hb.tets, dtype=np.int32)
    v = np.ascontiguousarray(hb.vert_to_tet, dtype=np.int32)
    b = Block()
    b.gid = int(getattr(hb, "gid", 0))
    b.num_orig_particles = int(hb.num_orig)
    b.num_particles = p.shape[0]
    b.particles = p.ctypes.data
    b.num_tets = t.shape[0]
    b.tets = t.ctypes.data
    b.vert_to_tet = v.ctypes.data
    b.bounds_min[:] = [float(x) for x in np.asarray(hb.bounds_min, np.float32)]
    b.bounds_max[:] = [float(x) for x in np.asarray(hb.bounds_max, np.float32)]
    keep = [p, t, v]
    gg, gc = getattr(hb, "ghost_gid", None), getattr(hb, "ghost_cell", None)
    if gg is not None and gc is not None:
        gg = np.ascontiguousarray(gg, dtype=np.int32)
        gc = np.ascontiguousarray(gc, dtype=np.int32)
        if gg.shape != (p.shape[0] - int(hb.num_orig),) or gc.shape != gg.shape:
            raise ValueError("ghost_gid / ghost_cell need one entry per ghost")
        b.ghost_gid = gg.ctypes.data
        b.ghost_cell = gc.ctypes.data
        keep += [gg, gc]
    return b, keep
