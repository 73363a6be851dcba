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
    # kernel's multi-word verdict masks) and beyond (line windows)
    out["sparse8"] = _uniform_blocks(2, 150, 53, 0.2)
    # fewer still: boxes of ~10^5 points with lines of 17-64 points (line
    # windows) and one longer line (face-parallel probes) at 200^3
    out["coarse8"] = _uniform_blocks(2, 40, 54, 0.3)
    pts, bid, counts, bounds = synthetic.plummer_blocks(2, 500, 4000, seed=52)
    out["plummer8"] = ingest.decomposed_blocks(pts, bid, bounds, 0.05)
    return out


def _bits(a):
    a = np.ascontiguousarray(a, np.float32).ravel()
    u = a.view(np.uint32).copy()
    u[np.isnan(a)] = 0x7FC00000
    return u


@pytest.mark.parametrize("name,g", [("uniform8", 33), ("uniform8", 40), ("plummer8", 33),
                                    ("sparse8", 48), ("sparse8", 72), ("coarse8", 200)])
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


@pytest.mark.parametrize("shape,cic", [((2, 2, 2), False), ((2, 2, 3), False),
                                       ((2, 2, 3), True)])
def test_multiblock_projection_and_write(oracle, shape, cic, tmp_path):
    """2-D projection over z-stacked blocks: every block's [y][x] plane
    bit-exact against the oracle's model (own points by x-y, routed points
    by their 3-D index), then WriteGrid's file (ProjectGrid's float adds into
    the z = 0 planes in gid order, the z = 0 planes at their offsets)
    byte-identical to the oracle's."""
    from .test_grid_io import shaped_blocks
    say("multiblock projection", shape, "cic" if cic else "tess")
    blocks = shaped_blocks(shape, 400, 71 + shape[2], 0.15)
    g = (33, 33, 33)
    alg = D.DENSE_CIC if cic else D.DENSE_TESS
    res = D.dense(alg, 0, None, None, True, (0.0, 0.0, 1.0), 1.0, 1e-4, g, blocks)
    want = oracle.dense_multi([oracle.Block.from_host(b) for b in blocks], g, cic=cic,
                              project=True)
    for i in range(len(blocks)):
        nm = res.block_num_idx[i]
        assert res.density[i].shape == (nm[1], nm[0])
        assert np.array_equal(_bits(res.density[i]), _bits(want["density"][i])), i
    assert res.stats["points_offblock"] == want["dropped"]
    gids = [b.gid for b in blocks]
    path = tmp_path / "plane.raw"
    D.write_blocks(path, gids, [np.ascontiguousarray(d.ravel()) for d in res.density],
                   res.block_min_idx, res.block_num_idx, g, project=True)
    assert path.read_bytes() == oracle.global_grid(want, g, project=True, gids=gids).tobytes()


def test_multiblock_write_3d(oracle, multiblocks, tmp_path):
    """WriteGrid's 3-D file of a multi-block run: every block's grid at its
    offsets, byte-identical to the oracle's assembly."""
    blocks = multiblocks["uniform8"]
    say("multiblock write")
    g = (33, 33, 33)
    res = D.dense(D.DENSE_TESS, 0, None, None, False, None, 1.0, 1e-4, g, blocks)
    want = oracle.dense_multi([oracle.Block.from_host(b) for b in blocks], g)
    path = tmp_path / "grid.raw"
    D.write_blocks(path, [b.gid for b in blocks],
                   [np.ascontiguousarray(d.ravel()) for d in res.density],
                   res.block_min_idx, res.block_num_idx, g)
    assert path.read_bytes() == oracle.global_grid(want, g).tobytes()
