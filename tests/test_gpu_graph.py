"""Graph replay and the work-split autotuner, through the C ABI.

With the graph option on, the first run records every launch of the path
(fills, circumcenters, cell kernels, fold, statistics, counter copies) into a
HIP graph and later runs replay it.  Each replay must give exactly the eager
result, also with other work and synchronisations between replays, and a
change of arguments must re-record.  Bar: bit-exact densities (NaN
canonicalised), identical statistics."""
import sys

import numpy as np
import pytest

from tess2_amd import host as D
from tess2_amd import ingest, synthetic

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _bits(a):
    a = np.ascontiguousarray(a, np.float32).ravel()
    u = a.view(np.uint32).copy()
    u[np.isnan(a)] = 0x7FC00000
    return u


@pytest.fixture(scope="module")
def single():
    return ingest.periodic_block(synthetic.uniform(4096, seed=61), 4.0 / 16)


@pytest.fixture(scope="module")
def multi():
    pts, bid, counts, bounds = synthetic.plummer_blocks(2, 500, 3000, seed=62)
    return ingest.decomposed_blocks(pts, bid, bounds, 0.05)


def _results(run, nb, exact=False):
    out = []
    for b in range(nb):
        r = [_bits(run.fetch(b)["density"])]
        if exact:
            r.append(run.owner(b))
        out.append(r)
    return out


def _same(a, b):
    assert len(a) == len(b)
    for x, y in zip(a, b):
        for u, v in zip(x, y):
            assert np.array_equal(u, v)


STAT_KEYS = ("cells_scan", "cells_cic", "cells_incomplete", "cells_outside",
             "cells_degenerate", "points_deposited", "points_offblock", "points_routed",
             "points_shared", "fp64_fallbacks", "tot_mass", "max_dense")


def _replays(blocks, kw, exact=False, n=4):
    nb = len(blocks)
    eager = D.StagedRun(blocks)
    eager.run(**kw)
    st0 = eager.sync()
    want = _results(eager, nb, exact)
    run = D.StagedRun(blocks)
    run.set_option("graph", 1)
    for i in range(n):
        run.run(**kw)
        st = run.sync()
        for k in STAT_KEYS:
            assert st[k] == st0[k], (i, k, st[k], st0[k])
        _same(_results(run, nb, exact), want)
        # unrelated eager work and a synchronisation between replays
        eager.run(**kw)
        eager.sync()
    return run, want


def test_graph_replay_single_block(oracle, single):
    say("graph single")
    kw = dict(glo_num_idx=(32, 32, 32))
    run, want = _replays([single], kw)
    ref = oracle.dense_single(oracle.Block.from_host(single), (32, 32, 32))["density"]
    assert np.array_equal(want[0][0], _bits(ref))


def test_graph_rerecords_on_new_arguments(oracle, single):
    say("graph new args")
    run = D.StagedRun([single])
    run.set_option("graph", 1)
    for g in (24, 32, 24, 20):
        run.run((g, g, g))
        run.sync()
        ref = oracle.dense_single(oracle.Block.from_host(single), (g, g, g))["density"]
        assert np.array_equal(_bits(run.fetch(0)["density"]), _bits(ref)), g
    run.run((20, 20, 20), alg=D.DENSE_CIC)
    run.sync()
    ref = oracle.dense_single(oracle.Block.from_host(single), (20, 20, 20), cic=True)
    assert np.array_equal(_bits(run.fetch(0)["density"]), _bits(ref["density"]))


def test_graph_replay_exact(oracle, single):
    """Replays of the exact assignment equal the eager run, and that run the
    FP64 oracle: every owner, densities within the exact tests' 1e-5."""
    say("graph exact")
    kw = dict(glo_num_idx=(30, 30, 30), assignment=D.EXACT_VORONOI)
    run, want = _replays([single], kw, exact=True)
    ref = oracle.exact_dense(oracle.Block.from_host(single), (30, 30, 30))
    assert np.array_equal(want[0][1], ref["owner"])
    den = want[0][0].view(np.float32).astype(np.float64)
    w = ref["density"]
    assert np.array_equal(den == 0, w == 0)
    nz = w != 0
    assert (np.abs(den[nz] - w[nz]) / np.abs(w[nz])).max() <= 1e-5


def test_graph_replay_multiblock(oracle, multi):
    say("graph multiblock", [b.num_orig for b in multi])
    kw = dict(glo_num_idx=(33, 33, 33))
    run, want = _replays(multi, kw)
    ref = oracle.dense_multi([oracle.Block.from_host(b) for b in multi], (33, 33, 33))
    for i in range(len(multi)):
        assert np.array_equal(want[i][0], _bits(ref["density"][i])), i


def test_options_and_autotune_keep_results(oracle, multi):
    say("autotune multiblock")
    run = D.StagedRun(multi)
    assert run.get_option("graph") == 0
    run.set_option("chunk_small", 3)
    run.set_option("waves_per_cu", 2)
    assert run.get_option("chunk_small") == 3
    with pytest.raises(D.Tess2Error):
        run.set_option("chunk_large", 0)
    run.run((33, 33, 33))
    run.sync()
    ref = oracle.dense_multi([oracle.Block.from_host(b) for b in multi], (33, 33, 33))
    got = [_bits(run.fetch(i)["density"]) for i in range(len(multi))]
    for i in range(len(multi)):
        assert np.array_equal(got[i], _bits(ref["density"][i])), i
    run.set_option("graph", 1)
    ms, opts = run.autotune(reps=2)
    say("  best", ms, opts)
    assert ms > 0
    for i in range(len(multi)):
        assert np.array_equal(_bits(run.fetch(i)["density"]), got[i]), i


def test_graph_replay_exact_multiblock(oracle, multi):
    """Replays over blocks equal the eager run, whose owners (home keys) are
    the FP64 oracle's."""
    from .test_oracle_exact_multi import home_keys
    say("graph exact multiblock")
    kw = dict(glo_num_idx=(33, 33, 33), assignment=D.EXACT_VORONOI)
    run, want = _replays(multi, kw, exact=True)
    ref = oracle.exact_dense_multi([oracle.Block.from_host(b) for b in multi], home_keys(multi),
                                   (33, 33, 33))
    for k in range(len(multi)):
        assert np.array_equal(want[k][1], ref["owner"][k]), k
