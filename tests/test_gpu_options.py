"""dense() options against the oracle, through the C ABI: the 2-D x-y
projection (src/dense.cpp:1076-1089: every cell's points indexed by x and y
only, contributions of all z added), given grid bounds (GridStepParams with
num_given_bounds, src/dense.cpp:1742-1766), and non-default mass and eps.
Bar: bit-exact densities (NaN canonicalised), identical statistics of
deposited cells."""
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
def block():
    return ingest.periodic_block(synthetic.grf_lattice(16, 0.3, seed=91), 4.0 / 16)


@pytest.mark.parametrize("cic", [False, True])
@pytest.mark.parametrize("g", [24, 33])
def test_projection_bitexact(oracle, block, g, cic):
    say("projection", g, "cic" if cic else "tess")
    alg = D.DENSE_CIC if cic else D.DENSE_TESS
    res = D.dense(alg, 0, None, None, True, (0.0, 0.0, 1.0), 1.0, 1e-4, (g, g, g), [block])
    assert res.density[0].shape == (g, g)
    want = oracle.dense_single(oracle.Block.from_host(block), (g, g, g), cic=cic, project=True)
    assert np.array_equal(_bits(res.density[0]), _bits(want["density"]))


@pytest.mark.parametrize("nb", [1, 2, 3])
def test_given_bounds_bitexact(oracle, block, nb):
    say("given bounds", nb)
    mins, maxs = (-0.1, 0.05, -0.2), (1.2, 0.95, 1.1)
    g = (30, 30, 30)
    res = D.dense(D.DENSE_TESS, nb, mins, maxs, False, None, 1.0, 1e-4, g, [block])
    want = oracle.dense_single(oracle.Block.from_host(block), g, num_given=nb,
                               given_mins=np.array(mins, np.float32),
                               given_maxs=np.array(maxs, np.float32))
    assert np.array_equal(res.grid_step_size, want["grid_step_size"])
    assert np.array_equal(res.grid_phys_mins, want["grid_phys_mins"])
    assert np.array_equal(_bits(res.density[0]), _bits(want["density"]))


@pytest.mark.parametrize("mass,eps", [(2.5, 1e-4), (1.0, 3e-3), (0.125, 0.0)])
def test_mass_and_eps_bitexact(oracle, block, mass, eps):
    say("mass/eps", mass, eps)
    g = (28, 28, 28)
    res = D.dense(D.DENSE_TESS, 0, None, None, False, None, mass, eps, g, [block],
                  want_cells=True)
    want = oracle.dense_single(oracle.Block.from_host(block), g, mass=mass, eps=eps)
    assert np.array_equal(res.cells[0][0], want["status"])
    assert np.array_equal(res.cells[0][1], want["npts"])
    assert np.array_equal(_bits(res.density[0]), _bits(want["density"]))
