"""Host-side mirror of the reference's interface for the density path.

Same names, argument meaning and outputs as the reference, over the HIP
library (tess2_amd/libtess2_amd.so):

  fill_circumcenters(tets, particles)       src/volume.cpp:6-11
  volume(v, block) / volumes(block)         src/volume.cpp:13-54
  grid_step_params(...)                     src/dense.cpp:1712-1767
  block_grid_params(...)                    src/dense.cpp:575-648
  dense(alg, ..., blocks)                   src/dense.cpp:30-103
  write_grid(path, density, glo_num_idx)    src/dense.cpp:751-870 (raw output)
  project_grid(gids, density, ...)          src/dense.cpp:881-1023
  write_blocks(path, gids, density, ...)    src/dense.cpp:751-1023 (block subarrays)
  ghost_sends(block, boxes, first)          src/tess.cpp:492-628 (incomplete_cells test)

Errors raise Tess2Error (the reference asserts or aborts instead).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _native as N
from ._native import Tess2Error  # noqa: F401

DENSE_TESS = N.TESS2_DENSE_TESS
DENSE_CIC = N.TESS2_DENSE_CIC
REFERENCE_SCAN = N.ASSIGN_REFERENCE_SCAN
EXACT_VORONOI = N.ASSIGN_EXACT_VORONOI
MIXED = N.PREC_MIXED
FP64 = N.PREC_FP64

_ctx = None


def context() -> N.Context:
    global _ctx
    if _ctx is None or _ctx.h is None:
        _ctx = N.Context(0)
    return _ctx


def fill_circumcenters(tets: np.ndarray, particles: np.ndarray, ctx=None) -> np.ndarray:
    """Circumcenter of every tet, float32 (T, 3)."""
    ctx = ctx or context()
    t = np.ascontiguousarray(tets, dtype=np.int32).reshape(-1, 8)
    p = np.ascontiguousarray(particles, dtype=np.float32).reshape(-1, 3)
    out = np.zeros((t.shape[0], 3), np.float32)
    ctx.check(N.lib().tess2_fill_circumcenters(ctx.h, t.ctypes.data, t.shape[0], p.ctypes.data,
                                               p.shape[0], out.ctypes.data),
              "fill_circumcenters")
    return out


def volumes(block, ctx=None) -> np.ndarray:
    """volume() of every original cell; -1 for infinite cells."""
    ctx = ctx or context()
    b, keep = N.make_block(block)
    out = np.zeros(block.num_orig, np.float32)
    ctx.check(N.lib().tess2_volumes(ctx.h, C.byref(b), out.ctypes.data), "volumes")
    return out


def volume(v: int, block, ctx=None) -> float:
    """volume() of one cell (computes all cells of the block)."""
    return float(volumes(block, ctx)[v])


def grid_step_params(num_given_bounds, given_mins, given_maxs, data_mins, data_maxs,
                     glo_num_idx):
    """-> (grid_phys_mins, grid_phys_maxs, grid_step_size)"""
    gmn, gmx, st = N.f3(None), N.f3(None), N.f3(None)
    N.lib().tess2_grid_step_params(int(num_given_bounds), N.f3(given_mins or (0, 0, 0)),
                                   N.f3(given_maxs or (0, 0, 0)), N.f3(data_mins),
                                   N.f3(data_maxs), gmn, gmx, st, N.i3(glo_num_idx))
    return (np.array(gmn[:], np.float32), np.array(gmx[:], np.float32),
            np.array(st[:], np.float32))


def block_grid_params(bounds_min, bounds_max, grid_phys_mins, grid_step_size, eps, data_mins,
                      data_maxs, glo_num_idx):
    """-> (block_min_idx, block_max_idx, block_num_idx)"""
    mn, mx, nm = N.i3((0, 0, 0)), N.i3((0, 0, 0)), N.i3((0, 0, 0))
    N.lib().tess2_block_grid_params(N.f3(bounds_min), N.f3(bounds_max), mn, mx, nm,
                                    N.f3(grid_phys_mins), N.f3(grid_step_size),
                                    float(np.float32(eps)), N.f3(data_mins), N.f3(data_maxs),
                                    N.i3(glo_num_idx))
    return (np.array(mn[:], np.int32), np.array(mx[:], np.int32), np.array(nm[:], np.int32))


@dataclass
class DenseResult:
    density: list            # per block: float32 array (z, y, x) or (y, x) when projecting
    data_mins: np.ndarray
    data_maxs: np.ndarray
    grid_phys_mins: np.ndarray
    grid_phys_maxs: np.ndarray
    grid_step_size: np.ndarray
    block_num_idx: list
    block_min_idx: list
    stats: dict


def _block_grids(blocks, num_given, given_mins, given_maxs, glo_num_idx, eps):
    dmin = np.min([np.asarray(b.bounds_min, np.float32) for b in blocks], axis=0)
    dmax = np.max([np.asarray(b.bounds_max, np.float32) for b in blocks], axis=0)
    gmin, gmax, step = grid_step_params(num_given, given_mins, given_maxs, dmin, dmax,
                                        glo_num_idx)
    grids = [block_grid_params(b.bounds_min, b.bounds_max, gmin, step, eps, dmin, dmax,
                               glo_num_idx) for b in blocks]
    return dmin, dmax, gmin, gmax, step, grids


def dense(alg, num_given_bounds, given_mins, given_maxs, project, proj_plane, mass, eps,
          glo_num_idx, blocks, assignment=REFERENCE_SCAN, precision=MIXED, ctx=None,
          want_cells=False):
    """dense() over host blocks: density of every block on its block grid.
    Results come back in the order of `blocks`; the library sees them in
    ascending gid order (the order received points are added in)."""
    ctx = ctx or context()
    blocks = list(blocks)
    perm = sorted(range(len(blocks)), key=lambda i: getattr(blocks[i], "gid", 0))
    if perm != list(range(len(blocks))):
        res = dense(alg, num_given_bounds, given_mins, given_maxs, project, proj_plane, mass,
                    eps, glo_num_idx, [blocks[i] for i in perm], assignment, precision, ctx,
                    want_cells)
        inv = [0] * len(perm)
        for k, i in enumerate(perm):
            inv[i] = k
        res.density = [res.density[inv[i]] for i in range(len(perm))]
        res.block_num_idx = [res.block_num_idx[inv[i]] for i in range(len(perm))]
        res.block_min_idx = [res.block_min_idx[inv[i]] for i in range(len(perm))]
        if want_cells:
            res.cells = [res.cells[inv[i]] for i in range(len(perm))]
        return res
    nb = len(blocks)
    cs = (N.Block * nb)()
    keep = []
    for i, hb in enumerate(blocks):
        b, k = N.make_block(hb)
        cs[i] = b
        keep.append(k)
    _, _, _, _, _, grids = _block_grids(blocks, num_given_bounds, given_mins, given_maxs,
                                        glo_num_idx, eps)
    outs, ptrs = [], (N.F32P * nb)()
    for i, (mn, mx, num) in enumerate(grids):
        shape = (int(num[1]), int(num[0])) if project else (int(num[2]), int(num[1]),
                                                            int(num[0]))
        a = np.zeros(shape, np.float32)
        outs.append(a)
        ptrs[i] = N.fptr(a)
    dmin, dmax, gmn, gmx, st = (N.f3(None) for _ in range(5))
    stats = N.Stats()
    r = N.lib().tess2_dense(ctx.h, int(alg), int(num_given_bounds), N.f3(given_mins),
                            N.f3(given_maxs), int(bool(project)), N.f3(proj_plane or (0, 0, 1)),
                            float(np.float32(mass)), dmin, dmax, gmn, gmx, st,
                            float(np.float32(eps)), N.i3(glo_num_idx), cs, nb, int(assignment),
                            int(precision), ptrs, C.byref(stats))
    ctx.check(r, "dense")
    res = DenseResult(density=outs, data_mins=np.array(dmin[:], np.float32),
                      data_maxs=np.array(dmax[:], np.float32),
                      grid_phys_mins=np.array(gmn[:], np.float32),
                      grid_phys_maxs=np.array(gmx[:], np.float32),
                      grid_step_size=np.array(st[:], np.float32),
                      block_num_idx=[g[2] for g in grids], block_min_idx=[g[0] for g in grids],
                      stats=stats.as_dict())
    if want_cells:
        res.cells = []
        for i, hb in enumerate(blocks):
            s = np.zeros(hb.num_orig, np.int32)
            n = np.zeros(hb.num_orig, np.int32)
            ctx.check(N.lib().tess2_dense_cells(ctx.h, i, s.ctypes.data, n.ctypes.data),
                      "dense_cells")
            res.cells.append((s, n))
    return res


def write_grid(path, density, glo_num_idx, project=False):
    """Raw float32, native endian, x fastest, truncating the file first."""
    d = np.ascontiguousarray(density, dtype=np.float32)
    r = N.lib().tess2_write_grid(str(path).encode(), d.ctypes.data, N.i3(glo_num_idx),
                                 int(bool(project)))
    if r != 0:
        raise Tess2Error(f"write_grid: {N.ERRORS.get(r, r)}")


def _block_arrays(gids, density, block_min_idx, block_num_idx):
    n = len(gids)
    if not (len(density) == len(block_min_idx) == len(block_num_idx) == n):
        raise ValueError("one density, block_min_idx and block_num_idx per gid")
    for d in density:
        if not (isinstance(d, np.ndarray) and d.dtype == np.float32 and d.flags.c_contiguous):
            raise TypeError("block densities must be C-contiguous float32 arrays "
                            "(they are updated in place)")
    g = np.ascontiguousarray(gids, np.int32)
    mn = np.ascontiguousarray(np.asarray(block_min_idx).reshape(n, 3), np.int32)
    nm = np.ascontiguousarray(np.asarray(block_num_idx).reshape(n, 3), np.int32)
    ptrs = (N.F32P * n)()
    for i, d in enumerate(density):
        ptrs[i] = N.fptr(d)
    return n, g, mn, nm, ptrs


def project_grid(gids, density, block_min_idx, block_num_idx):
    """ProjectGrid(): adds the [y][x] planes of blocks above z index 0 into the
    plane of the z = 0 block of their x-y column, in ascending gid order; the
    root planes (arrays in `density`) are updated in place."""
    n, g, mn, nm, ptrs = _block_arrays(gids, density, block_min_idx, block_num_idx)
    r = N.lib().tess2_project_grid(n, g.ctypes.data, ptrs, mn.ctypes.data, nm.ctypes.data)
    if r != 0:
        raise Tess2Error(f"project_grid: {N.ERRORS.get(r, r)}")


def write_blocks(path, gids, density, block_min_idx, block_num_idx, glo_num_idx,
                 project=False):
    """WriteGrid() over blocks: each block's grid at its place in the global
    raw float32 array (file truncated first); with project, ProjectGrid first
    (root planes updated in place) and the z = 0 blocks' planes written."""
    n, g, mn, nm, ptrs = _block_arrays(gids, density, block_min_idx, block_num_idx)
    r = N.lib().tess2_write_blocks(str(path).encode(), n, g.ctypes.data, ptrs, mn.ctypes.data,
                                   nm.ctypes.data, N.i3(glo_num_idx), int(bool(project)))
    if r != 0:
        raise Tess2Error(f"write_blocks: {N.ERRORS.get(r, r)}")


def ghost_sends(block, boxes, first=0, ctx=None) -> np.ndarray:
    """The test of incomplete_cells() (src/tess.cpp:492-628) on the GPU: for
    every original particle of the block, a bit mask of the neighbour boxes
    (n x 6: min xyz, max xyz, wrapped; n <= 64) it must be sent to;
    neighbours before `first` are not tested."""
    ctx = ctx or context()
    b, keep = N.make_block(block)
    bx = np.ascontiguousarray(boxes, dtype=np.float32).reshape(-1, 6)
    out = np.zeros(block.num_orig, np.uint64)
    ctx.check(N.lib().tess2_ghost_sends(ctx.h, C.byref(b), bx.shape[0], bx.ctypes.data,
                                        int(first), out.ctypes.data), "ghost_sends")
    return out


class StagedRun:
    """Blocks resident in HBM; each run() replays the whole path on the device.
    Blocks are kept in ascending gid order (fetch/cells index that order)."""

    def __init__(self, blocks, ctx=None):
        self.ctx = ctx or N.Context(0)
        self.blocks = sorted(blocks, key=lambda b: getattr(b, "gid", 0))
        nb = len(self.blocks)
        cs = (N.Block * nb)()
        keep = []
        for i, hb in enumerate(self.blocks):
            b, k = N.make_block(hb)
            cs[i] = b
            keep.append(k)
        self.ctx.check(N.lib().tess2_stage(self.ctx.h, cs, nb), "stage")

    def run(self, glo_num_idx, alg=DENSE_TESS, mass=1.0, eps=1e-4, project=False,
            num_given_bounds=0, given_mins=None, given_maxs=None, assignment=REFERENCE_SCAN,
            precision=MIXED, want_volumes=True):
        self.ctx.check(N.lib().tess2_run(self.ctx.h, int(alg), int(num_given_bounds),
                                         N.f3(given_mins), N.f3(given_maxs),
                                         int(bool(project)), N.f3((0, 0, 1)),
                                         float(np.float32(mass)), float(np.float32(eps)),
                                         N.i3(glo_num_idx), int(assignment), int(precision),
                                         int(bool(want_volumes))), "run")
        grids = _block_grids(self.blocks, num_given_bounds, given_mins, given_maxs,
                             glo_num_idx, np.float32(eps))[5]
        self.ngrid = [int(g[2][0]) * int(g[2][1]) * (1 if project else int(g[2][2]))
                      for g in grids]

    def _grid_size(self, b, ngrid):
        """Grid points of block b in the last run (the library copies that many)."""
        n = self.ngrid[b]
        if ngrid is not None and int(ngrid) != n:
            raise ValueError(f"block {b} has {n} grid points, not {ngrid}")
        return n

    def sync(self) -> dict:
        st = N.Stats()
        self.ctx.check(N.lib().tess2_sync(self.ctx.h, C.byref(st)), "sync")
        return st.as_dict()

    def fetch(self, b, ngrid=None, volumes=False, circumcenters=False):
        hb = self.blocks[b]
        out = {}
        d = np.zeros(self._grid_size(b, ngrid), np.float32) if ngrid is not False else None
        v = np.zeros(hb.num_orig, np.float32) if volumes else None
        c = np.zeros((hb.num_tets, 3), np.float32) if circumcenters else None
        self.ctx.check(N.lib().tess2_fetch(self.ctx.h, int(b),
                                           d.ctypes.data if d is not None else None,
                                           v.ctypes.data if v is not None else None,
                                           c.ctypes.data if c is not None else None), "fetch")
        out["density"], out["volumes"], out["circumcenters"] = d, v, c
        return out

    def owner(self, b, ngrid=None):
        o = np.zeros(self._grid_size(b, ngrid), np.int32)
        self.ctx.check(N.lib().tess2_fetch_owner(self.ctx.h, int(b), o.ctypes.data), "owner")
        return o

    def cells(self, b):
        hb = self.blocks[b]
        s = np.zeros(hb.num_orig, np.int32)
        n = np.zeros(hb.num_orig, np.int32)
        self.ctx.check(N.lib().tess2_dense_cells(self.ctx.h, int(b), s.ctypes.data,
                                                 n.ctypes.data), "cells")
        return s, n

    def kernel_time(self):
        ms = C.c_float()
        n = C.c_int()
        self.ctx.check(N.lib().tess2_kernel_time(self.ctx.h, C.byref(ms), C.byref(n)),
                       "kernel_time")
        return ms.value, n.value

    OPTIONS = {"graph": N.OPT_GRAPH, "waves_per_cu": N.OPT_WAVES_PER_CU,
               "chunk_small": N.OPT_CHUNK_SMALL, "chunk_medium": N.OPT_CHUNK_MEDIUM,
               "chunk_large": N.OPT_CHUNK_LARGE, "cell_kernel": N.OPT_CELL_KERNEL}

    def set_option(self, name, value):
        """Launch option (graph replay, persistent waves, chunk sizes)."""
        self.ctx.check(N.lib().tess2_set_option(self.ctx.h, self.OPTIONS[name], int(value)),
                       "set_option " + name)

    def get_option(self, name):
        v = C.c_int()
        self.ctx.check(N.lib().tess2_get_option(self.ctx.h, self.OPTIONS[name], C.byref(v)),
                       "get_option " + name)
        return v.value

    def autotune(self, reps=3):
        """Pick the fastest work split for the last run's arguments; returns
        the best median run time (ms) and the chosen options."""
        ms = C.c_float()
        self.ctx.check(N.lib().tess2_autotune(self.ctx.h, int(reps), C.byref(ms)), "autotune")
        return ms.value, {k: self.get_option(k) for k in self.OPTIONS}
