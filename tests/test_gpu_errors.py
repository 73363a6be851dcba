"""Error behaviour of the C ABI: invalid arguments come back as negative
status codes with a message (tess2_last_error), never as an abort, and the
context stays usable afterwards.  The reference asserts or calls MPI_Abort
on such input (src/dense.cpp:1027-1036); the library reports instead."""
import ctypes as C
import sys

import numpy as np
import pytest

from tess2_amd import _native as N
from tess2_amd import host as D
from tess2_amd import ingest, synthetic

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def block():
    return ingest.periodic_block(synthetic.uniform(1500, seed=81), 0.3)


def _expect(err, fn, *a, **k):
    with pytest.raises(D.Tess2Error) as e:
        fn(*a, **k)
    assert err in str(e.value), str(e.value)


def test_run_argument_errors(block):
    say("errors: run arguments")
    run = D.StagedRun([block])
    _expect("invalid argument", run.run, (1, 16, 16))                       # < 2 points per axis
    _expect("invalid argument", run.run, (16, 16, 16), alg=7)
    _expect("invalid argument", run.run, (16, 16, 16), assignment=5)
    _expect("invalid argument", run.run, (16, 16, 16), precision=3)
    _expect("invalid argument", run.run, (16, 16, 16), assignment=D.EXACT_VORONOI,
            project=True)
    # the context still works after every rejected call
    run.run((16, 16, 16))
    st = run.sync()
    assert st["cells_scan"] + st["cells_cic"] > 0


def test_projection_plane_must_be_xy(block):
    say("errors: projection plane")
    ctx = N.Context(0)
    b, keep = N.make_block(block)
    assert N.lib().tess2_stage(ctx.h, C.byref(b), 1) == 0
    r = N.lib().tess2_run(ctx.h, 0, 0, N.f3(None), N.f3(None), 1, N.f3((1, 0, 0)), 1.0,
                          1e-4, N.i3((16, 16, 16)), 0, 0, 1)
    assert r == -1
    assert b"x-y" in N.lib().tess2_last_error(ctx.h)
    ctx.close()


def test_stage_validation(block):
    say("errors: staging")
    ctx = N.Context(0)
    # a tet vertex out of range
    bad = ingest.HostBlock(particles=block.particles, tets=block.tets.copy(),
                           vert_to_tet=block.vert_to_tet, num_orig=block.num_orig,
                           bounds_min=block.bounds_min, bounds_max=block.bounds_max)
    bad.tets[3, 1] = block.num_particles + 5
    b, keep = N.make_block(bad)
    assert N.lib().tess2_stage(ctx.h, C.byref(b), 1) == -1
    assert b"out of range" in N.lib().tess2_last_error(ctx.h)
    # several blocks must come in ascending gid order
    b0, k0 = N.make_block(block)
    arr = (N.Block * 2)()
    arr[0] = b0
    arr[1] = b0
    arr[0].gid, arr[1].gid = 3, 1
    assert N.lib().tess2_stage(ctx.h, arr, 2) == -1
    assert b"ascending gid" in N.lib().tess2_last_error(ctx.h)
    # nothing staged: a run is refused
    assert N.lib().tess2_stage(ctx.h, arr, 0) == 0
    r = N.lib().tess2_run(ctx.h, 0, 0, N.f3(None), N.f3(None), 0, N.f3((0, 0, 1)), 1.0, 1e-4,
                          N.i3((8, 8, 8)), 0, 0, 1)
    assert r == -1
    ctx.close()


def test_options_and_fetch_bounds(block):
    say("errors: options, fetch")
    run = D.StagedRun([block])
    _expect("invalid argument", run.set_option, "waves_per_cu", 0)
    _expect("invalid argument", run.set_option, "graph", 2)
    with pytest.raises(D.Tess2Error):
        run.autotune()                        # needs a previous run
    run.run((12, 12, 12))
    run.sync()
    with pytest.raises(ValueError):
        run.fetch(0, ngrid=11 ** 3)           # wrong buffer size is refused on the host
    out = np.zeros(4, np.float32)
    assert N.lib().tess2_fetch(run.ctx.h, 5, out.ctypes.data, None, None) == -1
    _expect("invalid argument", run.owner, 0)  # no exact-assignment run yet
