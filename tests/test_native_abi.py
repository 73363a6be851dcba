"""The C-ABI library loads without a GPU and exports every declared symbol;
its host-only helpers match the oracle bit for bit."""
import ctypes
import os
import re

import numpy as np

from tess2_amd import _native as N
from tess2_amd import host as D

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "tess2_amd.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(tess2_[a-z_0-9]+)\s*\(", src)))


def test_exports_every_declared_symbol():
    lib = ctypes.CDLL(N.LIB_PATH)
    names = _declared()
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(N.SIGNATURES), set(names) ^ set(N.SIGNATURES)


def test_host_grid_helpers_match_oracle(oracle):
    rng = np.random.default_rng(4)
    for _ in range(200):
        lo = (rng.random(3) - 0.5).astype(np.float32)
        hi = (lo + rng.random(3) * 3 + 0.01).astype(np.float32)
        g = rng.integers(2, 300, 3)
        a = D.grid_step_params(0, None, None, lo, hi, g)
        b = oracle.grid_step_params(lo, hi, g)
        for x, y in zip(a, b):
            assert np.array_equal(x.view(np.uint32), y.view(np.uint32))
        bl = lo + (hi - lo) * np.float32(rng.random())
        bh = np.minimum(hi, bl + (hi - lo) * np.float32(0.5))
        c = D.block_grid_params(bl, bh, a[0], a[2], 1e-4, lo, hi, g)
        d = oracle.block_grid_params(bl, bh, b[0], b[2], 1e-4, lo, hi, g)
        for x, y in zip(c, d):
            assert np.array_equal(x, y)


def test_write_grid_raw_layout(tmp_path):
    d = np.arange(4 * 3 * 2, dtype=np.float32).reshape(2, 3, 4)
    f = tmp_path / "g.raw"
    f.write_bytes(b"x" * 1000)                     # truncated first
    D.write_grid(f, d, (4, 3, 2))
    raw = np.fromfile(f, dtype=np.float32)
    assert raw.size == 24 and np.array_equal(raw, d.ravel())
