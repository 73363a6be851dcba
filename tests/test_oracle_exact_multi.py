"""The FP64 oracle of the exact assignment over several blocks, pinned on
the CPU: with one block it is the single-block oracle (owner, status, count,
density identical); over eight blocks whose ghost layer covers every
nearest site, each in-block grid point's owner is the global brute-force
nearest original (numpy, ties to the lowest home key), every in-block grid
point is owned exactly once across the blocks, and the deposited mass is
the cells' mass."""
import numpy as np
import pytest

from tess2_amd import ingest, synthetic


def home_keys(blocks):
    """rank << 23 | cell of every particle (blocks in ascending gid = rank)."""
    out = []
    for k, b in enumerate(blocks):
        own = (np.int64(k) << 23) | np.arange(b.num_orig, dtype=np.int64)
        gh = (b.ghost_gid.astype(np.int64) << 23) | b.ghost_cell.astype(np.int64)
        out.append(np.concatenate([own, gh]))
    return out


def uniform_blocks(per_block, seed, ghost):
    rng = np.random.default_rng(seed)
    pts, bid, bounds = [], [], []
    for b in range(8):
        lo = np.array([b & 1, (b >> 1) & 1, b >> 2]) * 0.5
        pts.append((lo + rng.random((per_block, 3)) * 0.5).astype(np.float32))
        bid.append(np.full(per_block, b))
        bounds.append((lo.astype(np.float32), (lo + 0.5).astype(np.float32)))
    return ingest.decomposed_blocks(np.concatenate(pts), np.concatenate(bid), bounds, ghost)


def test_one_block_is_the_single_block_oracle(oracle):
    hb = ingest.periodic_block(synthetic.uniform(1500, seed=12), 4.0 / 11)
    ob = oracle.Block.from_host(hb)
    g = (30, 30, 30)
    one = oracle.exact_dense(ob, g)
    multi = oracle.exact_dense_multi([ob], [np.arange(hb.num_particles)], g)
    assert multi["cic"] == int((one["status"] == 1).sum())
    assert np.array_equal(multi["owner"][0], one["owner"])
    assert np.array_equal(multi["status"][0], one["status"])
    assert np.array_equal(multi["npts"][0], one["npts"])
    if multi["cic"] == 0:
        assert np.allclose(multi["density"][0], one["density"], rtol=1e-12, atol=0)


def test_eight_blocks_against_brute_force(oracle):
    blocks = uniform_blocks(150, 53, 0.2)
    ob = [oracle.Block.from_host(b) for b in blocks]
    keys = home_keys(blocks)
    g = (40, 40, 40)
    res = oracle.exact_dense_multi(ob, keys, g)
    allp = np.concatenate([b.particles[:b.num_orig] for b in blocks]).astype(np.float64)
    allk = np.concatenate([k[:b.num_orig] for k, b in zip(keys, blocks)])
    dmin = np.min([b.bounds_min for b in blocks], axis=0)
    dmax = np.max([b.bounds_max for b in blocks], axis=0)
    gmin, _, step = oracle.grid_step_params(dmin, dmax, g)
    seen = {}
    for b, own in zip(blocks, res["owner"]):
        mn, _, nm = oracle.block_grid_params(b.bounds_min, b.bounds_max, gmin, step, 1e-4,
                                             dmin, dmax, g)
        lz, ly, lx = np.meshgrid(*(np.arange(int(nm[i])) for i in (2, 1, 0)), indexing="ij")
        gi = np.stack([lx.ravel() + mn[0], ly.ravel() + mn[1], lz.ravel() + mn[2]], 1)
        pos = (gmin[None, :] + gi.astype(np.float32) * step[None, :]).astype(np.float32)
        inb = np.all((pos >= b.bounds_min) & (pos <= b.bounds_max), axis=1)
        assert np.array_equal(own >= 0, inb)
        q = pos[inb].astype(np.float64)
        d = ((q[:, None, :] - allp[None, :, :]) ** 2).sum(-1)
        best = d.min(1)
        # lowest key among the equidistant
        want = np.where(d == best[:, None], allk[None, :], np.iinfo(np.int64).max).min(1)
        assert np.array_equal(own[inb], want)
        for t in map(tuple, gi[inb]):
            seen[t] = seen.get(t, 0) + 1
    assert max(seen.values()) == 1          # no grid point is owned twice
    dep = sum(float(d.sum()) for d in res["density"])
    div = float(np.float32(step[0] * step[1] * step[2]))
    ncells = sum(int(((s == 0) & (n > 0)).sum()) for s, n in zip(res["status"], res["npts"]))
    assert dep * div == pytest.approx(ncells * 1.0, rel=1e-9)
