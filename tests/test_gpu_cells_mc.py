"""Every single-block cell kernel (option TESS2_OPT_CELL_KERNEL: 0 one cell
per wave; 1, 2 two cells per wave in the geometry phase, cells_mc.hip, 2 with
the multi-word verdict masks) against the oracle, through the C ABI:
statuses, point counts, volumes and densities bit-exact on the single-block
inputs of the parity tests, including the degenerate lattices and cells
beyond the two-cell kernel's star capacity (deferred to the large-capacity
pass).  The other tests run the default choice."""
import sys

import numpy as np
import pytest

from tess2_amd import host as D
from tess2_amd import ingest, synthetic

pytestmark = pytest.mark.gpu

VARIANTS = [0, 1, 2]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _bits(a):
    a = np.ascontiguousarray(a, np.float32).ravel()
    u = a.view(np.uint32).copy()
    u[np.isnan(a)] = 0x7FC00000
    return u


@pytest.fixture(scope="module")
def blocks():
    from .test_gpu_capacity import _hub_points
    out = {}
    out["periodic4k"] = ingest.periodic_block(synthetic.uniform(4096, seed=31), 4.0 / 16)
    rng = np.random.default_rng(1234)
    out["p2"] = ingest.make_block(rng.random((4096, 3)).astype(np.float32), 4096, (0, 0, 0),
                                  (1, 1, 1))
    out["grf05"] = ingest.periodic_block(synthetic.grf_lattice(24, 0.5, seed=3), 4.0 / 24)
    out["lattice"] = ingest.periodic_block(synthetic.jittered_lattice(8, 0.0, seed=4), 3.0 / 8)
    p64 = synthetic.jittered_lattice64(20, 1e-6, seed=4)
    out["jitter"] = ingest.periodic_block(synthetic.jittered_lattice(20, 1e-6, seed=4), 3.0 / 20,
                                          p64=synthetic.lattice_tiebreak(p64, 20))
    out["hub"] = ingest.periodic_block(_hub_points(48, 7), 0.2)
    return out


CASES = [("periodic4k", 32), ("periodic4k", 8), ("p2", 64), ("grf05", 48), ("lattice", 17),
         ("jitter", 80), ("hub", 48)]


@pytest.mark.parametrize("variant", VARIANTS)
@pytest.mark.parametrize("name,g", CASES)
def test_cells_mc_bitexact(oracle, blocks, name, g, variant):
    hb = blocks[name]
    say("cells_mc", variant, name, g)
    run = D.StagedRun([hb])
    run.set_option("cell_kernel", variant)
    run.run((g, g, g))
    st = run.sync()
    out = run.fetch(0, ngrid=g ** 3, volumes=True)
    s, npts = run.cells(0)
    ob = oracle.Block.from_host(hb)
    want = oracle.dense_single(ob, (g, g, g))
    assert np.array_equal(s, want["status"])
    assert np.array_equal(npts, want["npts"])
    assert np.array_equal(_bits(out["density"]), _bits(want["density"]))
    assert np.array_equal(_bits(out["volumes"]), _bits(oracle.volumes(ob)))
    if name == "hub":
        assert st["cells_deferred"] >= 1


_BATCHED = {}


def _batched_blocks(name):
    from .test_gpu_multiblock import _uniform_blocks
    if name not in _BATCHED:
        if name == "uniform8":
            _BATCHED[name] = _uniform_blocks(2, 1200, 51, 0.1)
        elif name == "sparse8":
            _BATCHED[name] = _uniform_blocks(2, 150, 53, 0.2)
        elif name == "coarse8":
            _BATCHED[name] = _uniform_blocks(2, 40, 54, 0.3)
        else:
            pts, bid, counts, bounds = synthetic.plummer_blocks(2, 500, 4000, seed=52)
            _BATCHED[name] = ingest.decomposed_blocks(pts, bid, bounds, 0.05)
    return _BATCHED[name]


@pytest.mark.parametrize("variant", [0, 1, 3, 4])
@pytest.mark.parametrize("name,g", [("uniform8", 48), ("sparse8", 48), ("sparse8", 72),
                                    ("plummer8", 48), ("coarse8", 200)])
def test_cells_mc_batched_bitexact(oracle, name, g, variant):
    """The multi-block cell kernels (persistent, 0: one cell per wave; 1:
    two cells per wave; 3, the default: one cell per wave with the star
    search's record loads issued before the table work; 4: two cells per
    wave, one workgroup per two cells, no work queue), against the oracle's
    multi-block model: per-block statuses, counts, densities.  The
    sparse cases put cell boxes past the 512-point masks onto the line
    windows: lines of up to 16 points at 72^3, of 17-64 points (and one
    longer, on face-parallel probes) in coarse8 at 200^3."""
    blocks = _batched_blocks(name)
    say("cells_mc batched", name, g, "variant", variant)
    g = (g, g, g)
    run = D.StagedRun(blocks)
    run.set_option("cell_kernel", variant)
    run.run(g)
    run.sync()
    if (name, g) not in _BATCHED:
        _BATCHED[name, g] = oracle.dense_multi([oracle.Block.from_host(b) for b in blocks], g)
    want = _BATCHED[name, g]
    for i in range(len(blocks)):
        s, npts = run.cells(i)
        assert np.array_equal(s, want["status"][i])
        assert np.array_equal(npts, want["npts"][i])
        assert np.array_equal(_bits(run.fetch(i)["density"]), _bits(want["density"][i])), i
