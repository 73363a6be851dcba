"""Host ingestion and the synthetic workloads (CPU only).

- The column-parallel Qhull (several processes, tets certified Delaunay for
  the whole set, neighbours rebuilt by face matching) must give every
  original particle the same star as one Qhull call on all points.
- DIY-style decomposition: each block holds its own particles first, in
  input order, then the other blocks' particles within the ghost width; its
  tessellation has valid neighbour links and vert_to_tet entries.
- The generators follow their stated distributions and seeds.
"""
import numpy as np

from tess2_amd import ingest, synthetic


def _stars(tets, n_orig):
    """For each of the first n_orig vertices, the sorted vertex 4-tuples of
    its incident tets."""
    out = [set() for _ in range(n_orig)]
    for t in tets[:, :4]:
        key = tuple(sorted(int(x) for x in t))
        for v in t:
            if v < n_orig:
                out[v].add(key)
    return out


def _check_links(tets):
    """tets[t][4+i] is the tet across the face opposite verts[i]: it shares
    exactly that face's three vertices, and links are mutual."""
    tv = tets[:, :4]
    for t in range(0, tets.shape[0], max(1, tets.shape[0] // 2000)):
        for i in range(4):
            nb = tets[t, 4 + i]
            if nb < 0:
                continue
            face = set(int(x) for k, x in enumerate(tv[t]) if k != i)
            assert face <= set(int(x) for x in tv[nb])
            assert t in set(int(x) for x in tets[nb, 4:])


def test_column_parallel_qhull_matches_single_call():
    p = synthetic.grf_lattice(24, 0.3, seed=11)            # 13824 clustered points
    g, _ = ingest.periodic_ghosts(p, 3.0 / 24)
    allp = np.concatenate([p, g]).astype(np.float32)
    n = p.shape[0]
    one = ingest.qhull_tets(allp)
    par = ingest.tessellate(allp, workers=4, columns=4, min_points=1000, closed_first=n)
    s1, s2 = _stars(one, n), _stars(par, n)
    assert s1 == s2
    _check_links(par)
    v2t = ingest.vert_to_tet(par, allp.shape[0])
    for v in range(0, n, 97):
        assert v in par[v2t[v], :4]


def test_jittered_lattice_columns_with_tiebreak_positions():
    """A lattice jittered below float32 resolution: the columns agree when the
    float64 tie-break positions decide the triangulation, the result equals
    one Qhull call on those positions, and it is a tessellation of the float32
    particles (every tet non-inverted in float64 tie-break space, every
    original star closed)."""
    n_side = 16
    p64 = synthetic.jittered_lattice64(n_side, 1e-6, seed=3)
    p = synthetic.jittered_lattice(n_side, 1e-6, seed=3)
    assert np.array_equal(p, p64.astype(np.float32))
    g, src = ingest.periodic_ghosts(p, 3.0 / n_side)
    allp = np.concatenate([p, g]).astype(np.float32)
    q64 = ingest.ghost_positions64(p, synthetic.lattice_tiebreak(p64, n_side), g, src)
    # the tie-break positions stay within 1e-2 spacing of the float32 particles
    assert np.abs(q64 - allp.astype(np.float64)).max() < 0.011 / n_side
    n = p.shape[0]
    one = ingest.qhull_tets(allp, q64)
    par = ingest.tessellate(allp, workers=4, columns=4, min_points=1000, closed_first=n,
                            qhull_points=q64)
    assert _stars(one, n) == _stars(par, n)
    _check_links(par)
    a, b, c, d = (q64[par[:, k]] for k in range(4))
    vol = np.einsum("ij,ij->i", b - a, np.cross(c - a, d - a))
    assert np.all(np.abs(vol) > 0)
    assert int(ingest.open_star_vertices(par, allp.shape[0])[:n].sum()) == 0


def test_decomposed_blocks_layout_and_links():
    pts, bid, counts, bounds = synthetic.plummer_blocks(2, 300, 1500, seed=7)
    w = 0.05
    blocks = ingest.decomposed_blocks(pts, bid, bounds, w)
    assert len(blocks) == 8
    for b, hb in enumerate(blocks):
        assert hb.gid == b
        own = pts[bid == b]
        assert hb.num_orig == own.shape[0] == counts[b]
        assert np.array_equal(hb.particles[:hb.num_orig], own)
        lo, hi = bounds[b]
        gh = hb.particles[hb.num_orig:]
        assert np.all(gh >= lo - np.float32(w)) and np.all(gh <= hi + np.float32(w))
        # every other block's particle within the ghost width is a ghost
        near = (bid != b) & np.all(pts >= lo - np.float32(w), 1) & np.all(pts <= hi + np.float32(w), 1)
        assert gh.shape[0] == int(near.sum())
        # each ghost's home: the block holding it as an original, and its index there
        assert hb.ghost_gid.shape == hb.ghost_cell.shape == (gh.shape[0],)
        for k in range(0, gh.shape[0], 7):
            home = blocks[int(hb.ghost_gid[k])]
            assert home.gid != b
            assert np.array_equal(home.particles[int(hb.ghost_cell[k])], gh[k])
            assert int(hb.ghost_cell[k]) < home.num_orig
        _check_links(hb.tets)
        v2t = hb.vert_to_tet
        for v in range(0, hb.num_particles, 37):
            if v2t[v] >= 0:
                assert v in hb.tets[v2t[v], :4]


def test_plummer_blocks_counts_and_bounds():
    pts, bid, counts, bounds = synthetic.plummer_blocks(4, 5000, 200000, seed=2)
    assert counts.shape == (64,)
    assert counts.min() >= 5000 and counts.max() <= 200000
    # log-uniform: the median of log-counts sits near the middle of the range
    mid = 0.5 * (np.log(5000) + np.log(200000))
    assert abs(np.median(np.log(counts)) - mid) < 0.6
    for b, (lo, hi) in enumerate(bounds[:8]):
        q = pts[bid == b]
        assert q.shape[0] == counts[b]
        assert np.all(q >= lo) and np.all(q < hi)
    pts2, _, counts2, _ = synthetic.plummer_blocks(4, 5000, 200000, seed=2)
    assert np.array_equal(counts, counts2) and np.array_equal(pts[:1000], pts2[:1000])


def test_grf_lattice_rms_and_wrap():
    n = 32
    p = synthetic.grf_lattice(n, 0.3, seed=3).astype(np.float64)
    assert p.shape == (n ** 3, 3)
    assert np.all(p >= 0.0) and np.all(p < 1.0)
    i = np.arange(n)
    lat = (np.stack(np.meshgrid(i, i, i, indexing="ij"), -1).reshape(-1, 3) + 0.5) / n
    d = p - lat
    d -= np.round(d)                                        # periodic displacement
    rms = np.sqrt(np.mean(np.sum(d * d, axis=1))) * n
    assert abs(rms - 0.3) < 0.03, rms
