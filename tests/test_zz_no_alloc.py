"""The allocation-free volume() entry point (tess2_volumes_noalloc) against
tess2_volumes, through the C ABI: device-resident inputs and output, the
library's fixed scratch array; two calls in a row, the second on a smaller
block, each bit-exact (NaN canonicalised) with the allocating entry point.
Sizes beyond the scratch and out-of-range indices are rejected before any
kernel reads the arrays."""
import ctypes as C
import sys

import numpy as np
import pytest

from tess2_amd import _native as N
from tess2_amd import host as D
from tess2_amd import ingest, synthetic

pytestmark = pytest.mark.gpu

E_ARG, E_CAPACITY = -1, -3      # TESS2_E_ARG, TESS2_E_CAPACITY


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _bits(a):
    a = np.ascontiguousarray(a, np.float32).ravel()
    u = a.view(np.uint32).copy()
    u[np.isnan(a)] = 0x7FC00000
    return u


class _Hip:
    """Device buffers through the HIP runtime the library itself links (a
    second runtime in the process, such as the one bundled with torch, would
    not share its device state)."""

    def __init__(self):
        N.lib()
        self.rt = C.CDLL("libamdhip64.so.7")
        self.bufs = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.rt.hipMalloc(C.byref(p), C.c_size_t(max(a.nbytes, 4))) == 0
        self.bufs.append(p)
        assert self.rt.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), 1) == 0
        return p.value

    def get(self, ptr, n, dtype):
        a = np.zeros(n, dtype)
        assert self.rt.hipMemcpy(a.ctypes.data_as(C.c_void_p), C.c_void_p(ptr),
                                 C.c_size_t(a.nbytes), 2) == 0
        return a

    def free(self):
        for p in self.bufs:
            self.rt.hipFree(p)
        self.bufs = []


def _device(hip, hb):
    p = hip.put(np.asarray(hb.particles, np.float32))
    t = hip.put(np.asarray(hb.tets, np.int32))
    v = hip.put(np.asarray(hb.vert_to_tet, np.int32))
    o = hip.put(np.full(hb.num_orig, -7.0, np.float32))
    return p, t, v, o


def _noalloc(ctx, hb, p, t, v, o, num_tets=None):
    return N.lib().tess2_volumes_noalloc(ctx.h, p, hb.num_particles, hb.num_orig, t,
                                         hb.num_tets if num_tets is None else num_tets, v, o)


def test_noalloc_volumes_match_tess2_volumes():
    big = ingest.periodic_block(synthetic.uniform(4096, seed=31), 4.0 / 16)
    small = ingest.periodic_block(synthetic.grf_lattice(12, 0.3, seed=5), 4.0 / 12)
    assert small.num_particles < big.num_particles and small.num_tets < big.num_tets
    ctx = N.Context(0)
    hip = _Hip()
    try:
        for name, hb in (("4096 uniform", big), ("12^3 grf", small)):
            say("noalloc volumes", name, hb.num_orig, hb.num_tets)
            want = D.volumes(hb)
            p, t, v, o = _device(hip, hb)
            r = _noalloc(ctx, hb, p, t, v, o)
            assert r == 0, N.lib().tess2_last_error(ctx.h)
            got = hip.get(o, hb.num_orig, np.float32)
            assert np.array_equal(_bits(got), _bits(want)), name
    finally:
        hip.free()


def test_noalloc_rejects_sizes_and_bad_indices():
    hb = ingest.periodic_block(synthetic.uniform(512, seed=3), 4.0 / 8)
    ctx = N.Context(0)
    hip = _Hip()
    try:
        p, t, v, o = _device(hip, hb)
        # 5M tets need more than the 64 MiB scratch: refused before any kernel
        say("noalloc capacity")
        assert _noalloc(ctx, hb, p, t, v, o, num_tets=5_000_000) == E_CAPACITY
        # a vertex id out of range: the device range check refuses the block
        say("noalloc range check")
        bad = hb.tets.copy()
        bad[3, 1] = hb.num_particles
        tb = hip.put(np.asarray(bad, np.int32))
        assert _noalloc(ctx, hb, p, tb, v, o) == E_ARG
        # and the next good call still works
        assert _noalloc(ctx, hb, p, t, v, o) == 0
        got = hip.get(o, hb.num_orig, np.float32)
        assert np.array_equal(_bits(got), _bits(D.volumes(hb)))
    finally:
        hip.free()
