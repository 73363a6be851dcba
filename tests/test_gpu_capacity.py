"""Cells beyond the fast kernel's capacities (64 star tets, 64 faces) go to
the large-capacity pass (512 tets, 128 faces); beyond those the run fails
with TESS2_E_CAPACITY instead of writing a wrong answer.  A "hub" input
forces both: one particle with n neighbours on a small sphere around it
(star ~2n tets, n faces) inside a uniform background.  Bar: bit-exact
statuses, counts, densities and volumes against the oracle, single-block
and batched."""
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


def _hub_points(nshell, seed, center=(0.5, 0.5, 0.5)):
    rng = np.random.default_rng(seed)
    c = np.array(center)
    bg = rng.random((3000, 3))
    bg = bg[np.linalg.norm(bg - c, axis=1) > 0.16]
    d = rng.normal(size=(nshell, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    return np.concatenate([c[None], c + 0.08 * d, bg]).astype(np.float32)


@pytest.mark.parametrize("nshell", [48, 80])   # star 92 / 156 tets, 48 / 80 faces
def test_large_cells_single_block(oracle, nshell):
    say("hub", nshell)
    hb = ingest.periodic_block(_hub_points(nshell, 7), 0.2)
    g = (48, 48, 48)
    res = D.dense(D.DENSE_TESS, 0, None, None, False, None, 1.0, 1e-4, g, [hb], want_cells=True)
    assert res.stats["cells_deferred"] >= 1
    want = oracle.dense_single(oracle.Block.from_host(hb), g)
    assert np.array_equal(res.cells[0][0], want["status"])
    assert np.array_equal(res.cells[0][1], want["npts"])
    assert np.array_equal(_bits(res.density[0]), _bits(want["density"]))
    assert np.array_equal(_bits(D.volumes(hb)), _bits(oracle.volumes(oracle.Block.from_host(hb))))


def test_large_cells_batched(oracle):
    say("hub, 8 blocks")
    p = _hub_points(80, 9, center=(0.25, 0.25, 0.25))
    ijk = np.minimum((p * 2).astype(int), 1)
    bid = ijk[:, 0] + 2 * ijk[:, 1] + 4 * ijk[:, 2]
    bounds = []
    for b in range(8):
        lo = np.array([b & 1, (b >> 1) & 1, b >> 2], np.float32) * 0.5
        bounds.append((lo, lo + np.float32(0.5)))
    blocks = ingest.decomposed_blocks(p, bid, bounds, 0.15)
    g = (40, 40, 40)
    res = D.dense(D.DENSE_TESS, 0, None, None, False, None, 1.0, 1e-4, g, blocks,
                  want_cells=True)
    assert res.stats["cells_deferred"] >= 1
    want = oracle.dense_multi([oracle.Block.from_host(b) for b in blocks], g)
    for i in range(len(blocks)):
        assert np.array_equal(res.cells[i][0], want["status"][i])
        assert np.array_equal(res.cells[i][1], want["npts"][i])
        assert np.array_equal(_bits(res.density[i]), _bits(want["density"][i]))


def _blob_points(n_blob, n_bg, seed):
    """A tight Gaussian blob in a uniform background."""
    rng = np.random.default_rng(seed)
    blob = 0.5 + 0.004 * rng.standard_normal((n_blob, 3))
    return np.concatenate([blob, rng.random((n_bg, 3))]).astype(np.float32)


@pytest.mark.parametrize("graph", [0, 1])
def test_shared_list_overflow_rerun(oracle, graph):
    """~72k blob cells on an 8^3 grid: each blob cell's box holds one grid
    point, (3, 3, 3), and a one-point-wide scan line always keeps its point
    (src/dense.cpp:1543-1544: min_xi = nx - 1 = 0 <= max_xi = 0), so every
    blob cell deposits there.  ~72k contributions beyond the second overflow
    the shared-point list (65536 entries at first); the run is redone with
    larger lists (in graph mode the graph is recorded again) and must come
    out bit-exact, with one segmented fold of ~72k contributions."""
    say("list overflow, graph", graph)
    hb = ingest.make_block(_blob_points(72000, 3000, 11), 75000, (0, 0, 0), (1, 1, 1))
    g = (8, 8, 8)
    want = oracle.dense_single(oracle.Block.from_host(hb), g)
    run = D.StagedRun([hb])
    run.set_option("graph", graph)
    for rep in range(2):
        run.run(g, want_volumes=False)
        st = run.sync()
        say("  run", rep, "points_shared", st["points_shared"])
        assert st["points_shared"] > 65536
        den = run.fetch(0, ngrid=512)["density"]
        s, npts = run.cells(0)
        assert np.array_equal(s, want["status"])
        assert np.array_equal(npts, want["npts"])
        assert np.array_equal(_bits(den), _bits(want["density"]))


def test_cell_beyond_all_capacities_is_an_error():
    say("hub 200")
    hb = ingest.periodic_block(_hub_points(200, 7), 0.2)   # 200 faces > 128
    with pytest.raises(D.Tess2Error) as e:
        D.dense(D.DENSE_TESS, 0, None, None, False, None, 1.0, 1e-4, (32, 32, 32), [hb])
    assert "capacit" in str(e.value)
