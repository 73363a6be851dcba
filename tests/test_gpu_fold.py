"""The ordered fold of shared grid points on heavy points, through the C ABI.

A dense Gaussian blob inside a sparse uniform background makes thousands of
cells far smaller than the grid spacing.  Their bounding boxes hold one grid
index, and the reference's scan reports a one-point run for such a box
whatever PtInCell says (its run bounds start at min = n-1 = 0, max = 0,
src/dense.cpp:1543-1544), so every blob cell deposits its whole mass on the
same grid point: one point collects ~20000 contributions -- beyond the
register class (16) and the one-wave LDS class (2048), into the chunk-sort +
merge class.  The reference adds them in cell order; bar: bit-exact
densities against the oracle (NaN canonicalised)."""
import sys

import numpy as np
import pytest

from tess2_amd import host as D
from tess2_amd import ingest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _bits(a):
    a = np.ascontiguousarray(a, np.float32).ravel()
    u = a.view(np.uint32).copy()
    u[np.isnan(a)] = 0x7FC00000
    return u


def _blob(n_blob, n_bg, sigma, seed):
    rng = np.random.default_rng(seed)
    c = np.array([0.5, 0.5, 0.5]) + 0.013          # away from the grid lines
    blob = c + sigma * rng.standard_normal((n_blob, 3))
    bg = rng.random((n_bg, 3))
    p = np.concatenate([blob, bg]) % 1.0
    return ingest.periodic_block(p.astype(np.float32), 0.25)


@pytest.fixture(scope="module")
def blob():
    return _blob(20000, 2000, 0.002, 71)


@pytest.mark.parametrize("g", [9, 17, 24])
def test_heavy_shared_points_bitexact(oracle, blob, g):
    say("fold blob", g)
    run = D.StagedRun([blob])
    run.run((g, g, g))
    st = run.sync()
    got = run.fetch(0)["density"]
    want = oracle.dense_single(oracle.Block.from_host(blob), (g, g, g))["density"]
    say("  shared contributions", st["points_shared"], "cic cells", st["cells_cic"])
    assert st["points_shared"] > 15000
    assert np.array_equal(_bits(got), _bits(want))
    # the same run again: the fold's scratch state starts over
    run.run((g, g, g))
    run.sync()
    assert np.array_equal(_bits(run.fetch(0)["density"]), _bits(want))


def test_heavy_shared_points_cic(oracle, blob):
    say("fold blob cic")
    run = D.StagedRun([blob])
    run.run((17, 17, 17), alg=D.DENSE_CIC)
    run.sync()
    want = oracle.dense_single(oracle.Block.from_host(blob), (17, 17, 17), cic=True)["density"]
    assert np.array_equal(_bits(run.fetch(0)["density"]), _bits(want))


@pytest.mark.parametrize("n_blob", [40, 63, 100, 600])
def test_mid_shared_points_bitexact(oracle, n_blob):
    """Smaller blobs put a few grid points in the one-wave class (17-2048
    contributions): about 40 and 63 land in the register sort (at most 64),
    100 and 600 in the LDS sort; both estimators."""
    blk = _blob(n_blob, 2000, 0.002, 72)
    g = (17, 17, 17)
    for alg, cic in ((D.DENSE_TESS, False), (D.DENSE_CIC, True)):
        say("fold mid blob", n_blob, "cic" if cic else "tess")
        run = D.StagedRun([blk])
        run.run(g, alg=alg)
        st = run.sync()
        assert st["points_shared"] >= n_blob - 2
        want = oracle.dense_single(oracle.Block.from_host(blk), g, cic=cic)["density"]
        assert np.array_equal(_bits(run.fetch(0)["density"]), _bits(want))
