"""GPU exact-Voronoi assignment against the FP64 oracle, through the C ABI.

Bar: owner of every grid point, per-cell status and point counts identical
(bit-exact assignment, ties to the lowest id); densities within a relative
error of 1e-5 of the FP64 oracle (max over nonzero points), zeros identical.
Both precisions: MIXED (FP32 with FP64 fallback) and FP64.
"""
import sys

import numpy as np
import pytest

from tess2_amd import host as D
from tess2_amd import ingest, synthetic

pytestmark = pytest.mark.gpu

REL_TOL = 1e-5


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def blocks():
    out = {}
    out["grf"] = ingest.periodic_block(synthetic.grf_lattice(24, 0.3, seed=5), 4.0 / 24)
    # rms 0.5 spacing, 2 grid points per spacing, as the 16.8M-particle workload
    out["grf05"] = ingest.periodic_block(synthetic.grf_lattice(24, 0.5, seed=3), 4.0 / 24)
    out["uniform"] = ingest.periodic_block(synthetic.uniform(3000, seed=41), 4.0 / 14)
    # exact dyadic lattice (sites at odd multiples of 1/16) on a 17^3 grid
    # (points at multiples of 1/16): many grid points exactly equidistant
    # from 2, 4 or 8 sites
    out["lattice"] = ingest.periodic_block(synthetic.jittered_lattice(8, 0.0, seed=4), 3.0 / 8)
    # lattice + 1e-6-spacing jitter (float32 mostly on the lattice) on a grid
    # of 4 points per spacing: grid points equidistant from lattice sites
    p64 = synthetic.jittered_lattice64(20, 1e-6, seed=4)
    out["jitter"] = ingest.periodic_block(synthetic.jittered_lattice(20, 1e-6, seed=4), 3.0 / 20,
                                          p64=synthetic.lattice_tiebreak(p64, 20))
    return out


def _check(oracle, hb, g, prec):
    run = D.StagedRun([hb])
    run.run((g, g, g), assignment=D.EXACT_VORONOI, precision=prec)
    st = run.sync()
    n = g ** 3
    own = run.owner(0, n)
    den = run.fetch(0, ngrid=n)["density"]
    s, npts = run.cells(0)
    want = oracle.exact_dense(oracle.Block.from_host(hb), (g, g, g))
    bad = np.nonzero(own != want["owner"])[0]
    assert bad.size == 0, (bad.size, bad[:5], own[bad[:5]], want["owner"][bad[:5]])
    assert np.array_equal(s, want["status"])
    assert np.array_equal(npts, want["npts"])
    w = want["density"]
    assert np.array_equal(den == 0, w == 0)
    nz = w != 0
    rel = np.abs(den[nz].astype(np.float64) - w[nz]) / np.abs(w[nz])
    assert rel.max() <= REL_TOL, rel.max()
    return st, rel


@pytest.mark.parametrize("prec", [D.MIXED, D.FP64])
@pytest.mark.parametrize("name,g", [("grf", 40), ("grf05", 48), ("uniform", 32), ("lattice", 17),
                                    ("jitter", 81)])
def test_exact_assignment(oracle, blocks, name, g, prec):
    say("exact", name, g, "mixed" if prec == D.MIXED else "fp64")
    st, rel = _check(oracle, blocks[name], g, prec)
    say("  fallbacks", st["fp64_fallbacks"], "max rel", float(rel.max()))
    if name == "lattice" and prec == D.MIXED:
        assert st["fp64_fallbacks"] > 0     # exact ties must go to FP64


def test_exact_then_reference_on_one_context(oracle, blocks):
    """Switching modes on staged blocks keeps both results exact."""
    hb = blocks["uniform"]
    run = D.StagedRun([hb])
    g = (28, 28, 28)
    run.run(g, assignment=D.EXACT_VORONOI)
    run.sync()
    run.run(g)
    run.sync()
    den = run.fetch(0, ngrid=28 ** 3)["density"]
    want = oracle.dense_single(oracle.Block.from_host(hb), g)["density"]
    assert np.array_equal(den.view(np.uint32), want.view(np.uint32))


@pytest.mark.parametrize("prec", [D.MIXED, D.FP64])
@pytest.mark.parametrize("per_block,ghost,g", [(150, 0.2, 48), (1200, 0.12, 64)])
def test_exact_assignment_over_blocks(oracle, per_block, ghost, g, prec):
    """Eight blocks with ghost homes: owners are home keys (rank << 23 |
    cell), counts are summed over every block that holds a cell's points."""
    from .test_oracle_exact_multi import home_keys, uniform_blocks
    say("exact 8 blocks", per_block, g, "mixed" if prec == D.MIXED else "fp64")
    blocks = uniform_blocks(per_block, 53 + per_block, ghost)
    run = D.StagedRun(blocks)
    run.run((g, g, g), assignment=D.EXACT_VORONOI, precision=prec)
    st = run.sync()
    want = oracle.exact_dense_multi([oracle.Block.from_host(b) for b in blocks],
                                    home_keys(blocks), (g, g, g))
    assert want["cic"] == 0 and st["cells_cic"] == 0
    worst = 0.0
    for k in range(len(blocks)):
        own = run.owner(k)
        bad = np.nonzero(own != want["owner"][k])[0]
        assert bad.size == 0, (k, bad.size, bad[:5], own[bad[:5]], want["owner"][k][bad[:5]])
        s, npts = run.cells(k)
        assert np.array_equal(s, want["status"][k])
        assert np.array_equal(npts, want["npts"][k])
        den = run.fetch(k)["density"]
        w = want["density"][k]
        assert np.array_equal(den == 0, w == 0)
        nz = w != 0
        if nz.any():
            worst = max(worst, float((np.abs(den[nz] - w[nz]) / np.abs(w[nz])).max()))
    assert worst <= REL_TOL, worst
    ncells = sum(int((s == 0).sum()) for s in want["status"])
    assert st["tot_mass"] == pytest.approx(ncells, rel=1e-5)
    say("  fallbacks", st["fp64_fallbacks"], "max rel", worst)


def test_exact_over_blocks_needs_ghost_homes():
    from .test_oracle_exact_multi import uniform_blocks
    blocks = uniform_blocks(150, 60, 0.2)
    for b in blocks:
        b.ghost_gid = b.ghost_cell = None
    run = D.StagedRun(blocks)
    with pytest.raises(D.Tess2Error) as e:
        run.run((24, 24, 24), assignment=D.EXACT_VORONOI)
    assert "ghost homes" in str(e.value)
    run.run((24, 24, 24))                  # the reference scan needs none
    run.sync()
