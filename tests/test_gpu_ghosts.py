"""Ghost selection on the GPU (tess2_ghost_sends, the test of
incomplete_cells(), src/tess.cpp:492-628) through the C ABI: every round of
the host's exchange protocol, on every block, the GPU's send masks equal the
oracle's bit for bit, and the ghosts they select close every original star
(each block's tessellation holds the Delaunay star of all points)."""
import sys

import numpy as np
import pytest

from tess2_amd import host as D
from tess2_amd import ingest

from .test_ghost_exchange import _input

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("name", ["plummer8", "uniform27"])
def test_gpu_ghost_sends_match_oracle_and_close_stars(oracle, name):
    say("ghost exchange", name)
    pts, bid, bounds = _input(name)
    calls = []

    def sel(hb, boxes, first):
        got = D.ghost_sends(hb, boxes, first)
        want = oracle.ghost_sends(oracle.Block.from_host(hb), boxes, first)
        calls.append(bool(np.array_equal(got, want)))
        return got

    blocks, rounds, sent = ingest.exchanged_blocks(pts, bid, bounds, sel)
    assert calls and all(calls), calls
    assert sent[-1] == 0 and sent[0] > 0
    res = ingest.stars_closed(blocks, pts, bid)
    assert res["closed"], res
