"""Pins the CPU oracle's tet geometry, stars, edge links and volumes.

What pins each restated function (the reference is not buildable in this
image: its sources include DIY headers that are absent):
  circumcenter (src/tet.cpp:37-66)   exact known answers on integer tets; FP64
                                     closed-form solve (equidistance) to a few
                                     float ulps on random tets; a wrong term,
                                     sign or operand order of the formula moves
                                     the result by O(1).
  neighbor_edges / complete          brute-force adjacency from the tet array;
    (src/tet.cpp:228-270,337-378)    first pair's tet is the start tet; each
                                     (u, ut) has ut containing both ends.
  fill_edge_link (src/tet.cpp:389)   ring == brute-force set of tets sharing the
                                     edge, consecutive ring tets share a face.
  volume (src/volume.cpp:13-54)      periodic sum == 1; per cell == convex hull
                                     volume of its Voronoi vertices (FP64).
  fill_vert_to_tet (src/tess.cpp:767) last (highest) tet containing the vertex.
BFS visit order itself is pinned by hand-worked known answers
(test_oracle_known_answers.py).
"""
import numpy as np
import pytest
from scipy.spatial import ConvexHull

from tess2_amd import ingest, synthetic


@pytest.fixture(scope="module")
def small_periodic():
    p = synthetic.uniform(2048, seed=11)
    return ingest.periodic_block(p, 4.0 / 2048 ** (1 / 3))


def _exact_cc(a, b, c, d):
    a, b, c, d = (np.asarray(x, np.float64) for x in (a, b, c, d))
    A = 2.0 * np.stack([a - d, b - d, c - d])
    rhs = np.array([a @ a - d @ d, b @ b - d @ d, c @ c - d @ d])
    return np.linalg.solve(A, rhs)


def test_circumcenter_known_answers(oracle):
    p = np.array([[1, 0, 0], [0, 1, 0], [0, 0, 1], [0, 0, 0],
                  [2, 0, 0], [0, 2, 0], [0, 0, 2]], np.float32)
    t = np.array([[0, 1, 2, 3, -1, -1, -1, -1],
                  [4, 5, 6, 3, -1, -1, -1, -1],
                  [3, 4, 5, 6, -1, -1, -1, -1]], np.int32)
    cc = oracle.circumcenters(t, p)
    assert np.array_equal(cc[0], np.float32([0.5, 0.5, 0.5]))
    assert np.array_equal(cc[1], np.float32([1, 1, 1]))
    assert np.array_equal(cc[2], np.float32([1, 1, 1]))


def test_circumcenter_closed_form(oracle):
    rng = np.random.default_rng(5)
    p = rng.random((4000, 3)).astype(np.float32)
    t = np.zeros((1000, 8), np.int32)
    t[:, :4] = np.arange(4000).reshape(1000, 4)
    cc = oracle.circumcenters(t, p).astype(np.float64)
    errs = []
    for i in range(1000):
        a, b, c, d = p[4 * i:4 * i + 4].astype(np.float64)
        vol = abs(np.linalg.det(np.stack([a - d, b - d, c - d]))) / 6
        e = _exact_cc(a, b, c, d)
        r = np.linalg.norm(e - d)
        # well-conditioned tets only: volume not tiny against r^3
        if vol < 1e-2 * r ** 3:
            continue
        errs.append(np.linalg.norm(cc[i] - e) / r)
    assert len(errs) > 300
    assert max(errs) < 1e-4


def test_vert_to_tet_last_wins(small_periodic):
    hb = small_periodic
    brute = np.full(hb.num_particles, -1)
    for t in range(hb.num_tets):
        for v in hb.tets[t, :4]:
            brute[v] = t
    assert np.array_equal(brute, hb.vert_to_tet)


def test_star_and_edge_links(oracle, small_periodic):
    hb = small_periodic
    tets = hb.tets
    rng = np.random.default_rng(0)
    cells = rng.choice(hb.num_orig, 60, replace=False)
    for v in cells:
        vt = hb.vert_to_tet[v]
        fin, nbrs = oracle.neighbor_edges(v, tets, vt)
        star = np.nonzero((tets[:, :4] == v).any(1))[0]
        brute_u = set(tets[star, :4].ravel().tolist()) - {int(v)}
        assert fin == 1 and oracle.complete(v, tets, vt) == 1
        assert set(u for u, _ in nbrs) == brute_u
        assert len(nbrs) == len(brute_u)
        assert nbrs[0][1] == vt
        for u, ut in nbrs:
            assert v in tets[ut, :4] and u in tets[ut, :4]
            ring = oracle.edge_link(v, u, ut, tets)
            both = set(np.nonzero((tets[:, :4] == v).any(1) & (tets[:, :4] == u).any(1))[0])
            assert set(ring) == both and len(ring) == len(both)
            assert ring[0] == ut
            for k in range(len(ring)):
                s0 = set(tets[ring[k], :4].tolist())
                s1 = set(tets[ring[(k + 1) % len(ring)], :4].tolist())
                assert len(s0 & s1) == 3


def test_complete_matches_hull_brute_force(oracle):
    p = synthetic.uniform(300, seed=3)
    hb = ingest.make_block(p, 300, (0, 0, 0), (1, 1, 1))
    for v in range(300):
        star = np.nonzero((hb.tets[:, :4] == v).any(1))[0]
        opp = []
        for t in star:
            for i in range(4):
                if hb.tets[t, i] != v:
                    opp.append(hb.tets[t, 4 + i])
        brute = int(-1 not in opp)
        assert oracle.complete(v, hb.tets, hb.vert_to_tet[v]) == brute


def test_volumes_periodic_sum(oracle, small_periodic):
    ob = oracle.Block.from_host(small_periodic)
    vol = oracle.volumes(ob)
    assert (vol > 0).all()
    assert abs(vol.astype(np.float64).sum() - 1.0) < 1e-5


def test_volume_equals_convex_hull(oracle, small_periodic):
    hb = small_periodic
    ob = oracle.Block.from_host(hb)
    cc = oracle.circumcenters(hb.tets, hb.particles)
    vol = oracle.volumes(ob, cc)
    for v in range(0, hb.num_orig, 97):
        star = np.nonzero((hb.tets[:, :4] == v).any(1))[0]
        hull = ConvexHull(cc[star].astype(np.float64)).volume
        assert abs(vol[v] - hull) <= 2e-4 * hull


def test_volume_infinite_cell(oracle):
    p = synthetic.uniform(200, seed=8)
    hb = ingest.make_block(p, 200, (0, 0, 0), (1, 1, 1))
    ob = oracle.Block.from_host(hb)
    vol = oracle.volumes(ob)
    for v in range(200):
        fin = oracle.complete(v, hb.tets, hb.vert_to_tet[v])
        assert (vol[v] == -1.0) == (fin == 0)
