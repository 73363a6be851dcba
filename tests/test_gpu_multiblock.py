"""Several DIY-style blocks in one batched launch, against the oracle's
multi-block model (own deposits in cell order, then points received from
other blocks by source gid and cell, routed to the block whose grid holds
their index).  Bar: bit-exact per-block densities, statuses and counts."""
import sys

import numpy as np
import pytest

from tess2_amd import host as D
from tess2_amd import ingest, synthetic

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _uniform_blocks(nside, per_block, seed, ghost):
    rng = np.random.default_rng(seed)
    w = 1.0 / nside
    pts, bid, bounds = [], [], []
    for b in range(nside ** 3):
        i, j, k = b % nside, (b // nside) % nside, b // nside ** 2
        lo = np.array([i, j, k]) * w
        p = (lo + rng.random((per_block, 3)) * w).astype(np.float32)
        pts.append(p)
        bid.append(np.full(per_block, b))
        bounds.append((lo.astype(np.float32), (lo + w).astype(np.float32)))
    return ingest.decomposed_blocks(np.concatenate(pts), np.concatenate(bid), bounds, ghost)


@pytest.fixture(scope="module")
def multiblocks():
    out = {}
    out["uniform8"] = _uniform_blocks(2, 1200, 51, 0.1)
    # few particles on a fine grid: cell boxes of ~100-500 points (the batched
    # kernel's multi-word verdict masks) and beyond (point-by-point probes)
    out["sparse8"] = _uniform_blocks(2, 150, 53, 0.2)
    pts, bid, counts, bounds = synthetic.plummer_blocks(2, 500, 4000, seed=52)
    out["plummer8"] = ingest.decomposed_blocks(pts, bid, bounds, 0.05)
    return out


def _bits(a):
    a = np.ascontiguousarray(a, np.float32).ravel()
    u = a.view(np.uint32).copy()
    u[np.isnan(a)] = 0x7FC00000
    return u


@pytest.mark.parametrize("name,g", [("uniform8", 33), ("uniform8", 40), ("plummer8", 33),
                                    ("sparse8", 48), ("sparse8", 72)])
def test_multiblock_reference_scan_bitexact(oracle, multiblocks, name, g):
    blocks = multiblocks[name]
    say("multiblock", name, g, [b.num_orig for b in blocks])
    res = D.dense(D.DENSE_TESS, 0, None, None, False, None, 1.0, 1e-4, (g, g, g), blocks,
                  want_cells=True)
    want = oracle.dense_multi([oracle.Block.from_host(b) for b in blocks], (g, g, g))
    for i in range(len(blocks)):
        assert np.array_equal(res.cells[i][0], want["status"][i])
        assert np.array_equal(res.cells[i][1], want["npts"][i])
        assert np.array_equal(_bits(res.density[i]), _bits(want["density"][i])), i
    assert res.stats["points_offblock"] == want["dropped"]


def test_multiblock_cic_bitexact(oracle, multiblocks):
    blocks = multiblocks["uniform8"]
    say("multiblock cic")
    res = D.dense(D.DENSE_CIC, 0, None, None, False, None, 1.0, 1e-4, (33, 33, 33), blocks)
    want = oracle.dense_multi([oracle.Block.from_host(b) for b in blocks], (33, 33, 33),
                              cic=True)
    for i in range(len(blocks)):
        assert np.array_equal(_bits(res.density[i]), _bits(want["density"][i])), i


def test_multiblock_order_independent_of_input_order(oracle, multiblocks):
    blocks = multiblocks["uniform8"]
    a = D.dense(D.DENSE_TESS, 0, None, None, False, None, 1.0, 1e-4, (33, 33, 33), blocks)
    rev = blocks[::-1]
    b = D.dense(D.DENSE_TESS, 0, None, None, False, None, 1.0, 1e-4, (33, 33, 33), rev)
    for i in range(len(blocks)):
        assert np.array_equal(_bits(a.density[i]), _bits(b.density[len(blocks) - 1 - i]))
