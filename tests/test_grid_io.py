"""WriteGrid() / ProjectGrid() over blocks (src/dense.cpp:751-1023) in the
C-ABI library, on the host (no GPU): the raw file of block subarrays against
numpy's assembly of the global array, and the 2-D projection's sum of
z-stacked block planes against the oracle's (float32 adds in ascending gid
order).  Also the oracle's multi-block projection against its own 3-D run:
a block's plane is h times the z-sum of its 3-D grid (to float rounding),
which fixes that routed points go to the block holding their 3-D index."""
import os

import numpy as np
import pytest

from tess2_amd import host as D
from tess2_amd import ingest


def _split(n, parts):
    cuts = np.linspace(0, n, parts + 1).astype(int)
    return [(int(a), int(b - a)) for a, b in zip(cuts[:-1], cuts[1:])]


def _fake_blocks(gnum, shape, seed, planes=False):
    """Block grids tiling a global grid of gnum points in shape blocks, gids
    shuffled; random float32 densities."""
    rng = np.random.default_rng(seed)
    sx, sy, sz = (_split(g, s) for g, s in zip(gnum, shape))
    mins, nums = [], []
    for k in range(shape[2]):
        for j in range(shape[1]):
            for i in range(shape[0]):
                mins.append((sx[i][0], sy[j][0], sz[k][0]))
                nums.append((sx[i][1], sy[j][1], sz[k][1]))
    n = len(mins)
    gids = rng.permutation(n).astype(np.int32)
    dens = [rng.random(nm[0] * nm[1] * (1 if planes else nm[2])).astype(np.float32) * 1e3
            for nm in nums]
    return gids, dens, np.array(mins, np.int32), np.array(nums, np.int32)


def test_write_blocks_3d(tmp_path):
    gnum = (10, 12, 14)
    gids, dens, mins, nums = _fake_blocks(gnum, (2, 3, 2), 1)
    path = tmp_path / "grid.raw"
    path.write_bytes(b"x" * 10_000_000)            # truncated first
    D.write_blocks(path, gids, dens, mins, nums, gnum)
    want = np.zeros(gnum[::-1], np.float32)
    for d, mn, nm in zip(dens, mins, nums):
        want[mn[2]:mn[2] + nm[2], mn[1]:mn[1] + nm[1], mn[0]:mn[0] + nm[0]] = \
            d.reshape(nm[2], nm[1], nm[0])
    assert path.read_bytes() == want.tobytes()


def test_project_grid_and_write(oracle, tmp_path):
    gnum = (9, 11, 13)
    gids, dens, mins, nums = _fake_blocks(gnum, (2, 2, 3), 2, planes=True)
    res = {"density": [d.copy() for d in dens],
           "grids": [(mn, mn + nm - 1, nm) for mn, nm in zip(mins, nums)]}
    want_planes = oracle.project_grid(res, gids)
    want_file = oracle.global_grid(res, gnum, project=True, gids=gids)
    # the three z levels make the order of the float adds visible
    z0 = [i for i in range(len(gids)) if mins[i][2] == 0]
    assert any(not np.array_equal(want_planes[i], dens[i]) for i in z0)
    mine = [d.copy() for d in dens]
    D.project_grid(gids, mine, mins, nums)
    for a, b in zip(mine, want_planes):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    path = tmp_path / "plane.raw"
    D.write_blocks(path, gids, [d.copy() for d in dens], mins, nums, gnum, project=True)
    assert path.read_bytes() == want_file.tobytes()
    assert want_file.size == gnum[0] * gnum[1]


def test_project_grid_needs_a_root():
    gids, dens, mins, nums = _fake_blocks((8, 8, 8), (1, 1, 2), 3, planes=True)
    with pytest.raises(D.Tess2Error):
        D.project_grid(gids[1:], dens[1:], mins[1:], nums[1:])


def shaped_blocks(shape, per_block, seed, ghost):
    """Uniform particles in a regular decomposition of the unit cube into
    shape[0] x shape[1] x shape[2] blocks, each tessellated with its ghosts."""
    rng = np.random.default_rng(seed)
    w = 1.0 / np.array(shape, np.float64)
    pts, bid, bounds = [], [], []
    for b in range(int(np.prod(shape))):
        ijk = np.array([b % shape[0], (b // shape[0]) % shape[1], b // (shape[0] * shape[1])])
        lo = ijk * w
        pts.append((lo + rng.random((per_block, 3)) * w).astype(np.float32))
        bid.append(np.full(per_block, b))
        bounds.append((lo.astype(np.float32), (lo + w).astype(np.float32)))
    return ingest.decomposed_blocks(np.concatenate(pts), np.concatenate(bid), bounds, ghost)


def test_oracle_projection_matches_3d_per_block(oracle):
    """Block planes against the same blocks' 3-D grids: grid 32^3 (points at
    k/31, never on the block faces at 1/2, 1/3, 2/3), so each grid point has
    one owning block in both runs."""
    blocks = shaped_blocks((2, 2, 3), 300, 61, 0.15)
    obs = [oracle.Block.from_host(b) for b in blocks]
    g = (32, 32, 32)
    r3 = oracle.dense_multi(obs, g)
    r2 = oracle.dense_multi(obs, g, project=True)
    h = np.float64(r3["grid_step_size"][2])
    assert r2["dropped"] == r3["dropped"]
    for k, (mn, _, nm) in enumerate(r3["grids"]):
        d3 = r3["density"][k].reshape(nm[2], nm[1], nm[0]).astype(np.float64)
        d2 = r2["density"][k].reshape(nm[1], nm[0]).astype(np.float64)
        want = d3.sum(0) * h
        assert np.allclose(d2, want, rtol=2e-5, atol=1e-3 * want.max()), k
    # after ProjectGrid the z = 0 planes hold the whole z column
    plane = oracle.global_grid(r2, g, project=True, gids=[b.gid for b in blocks])
    full = oracle.global_grid(r3, g).astype(np.float64).sum(0) * h
    assert np.allclose(plane, full, rtol=2e-5, atol=1e-3 * full.max())
