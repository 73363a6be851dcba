"""Pins the CPU oracle's density estimator (src/dense.cpp).

What pins each restated function:
  NewellNormal (:1131-1162)      == -normalize((p1-p0)x(p2-p0)) for triangles,
                                   area-weighted for planar polygons.
  PtInCell (:1172-1203)          hand-derived answers on a unit cube cell,
                                   including the eps band and sign rule.
  DistributeScalarCIC (:1787)    trilinear closed form; masses sum to the scalar;
                                   corner order z, y, x.
  GridStepParams (:1712-1767)    closed forms (cube, padded box, given bounds).
  BlockGridParams (:575-648)     hand-derived indices; shared-face blocks
                                   partition the global index range.
  IterateCells / CellGridPts /   mass conservation; scan points are a subset of
  CellInteriorGridPts             the brute-force PtInCell set; and counts
    (:220-313,1363-1701)          measured from the reference's scan built
                                   against stand-in headers, on a re-creatable
                                   input (tests/golden/reference_scan_counts.json,
                                   corroboration): deposited cells, scan misses,
                                   eps band and exactly-matching cells all
                                   reproduce.  The scan's state machine itself is
                                   pinned by test_oracle_known_answers.py.
  exact owner (FP64 oracle)      brute-force argmin with lowest-id ties.
"""
import json
import os

import numpy as np
import pytest

from tess2_amd import ingest, synthetic

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_newell_triangle_and_polygon(oracle):
    rng = np.random.default_rng(2)
    for _ in range(50):
        tri = rng.random((3, 3)).astype(np.float32)
        n = oracle.newell_normal(tri).astype(np.float64)
        t = tri.astype(np.float64)
        ref = np.cross(t[1] - t[0], t[2] - t[0])
        ref = -ref / np.linalg.norm(ref)
        assert np.allclose(n, ref, atol=2e-6)
    # planar square in z=0.5, counter-clockwise seen from +z -> normal -z
    sq = np.float32([[0, 0, .5], [1, 0, .5], [1, 1, .5], [0, 1, .5]])
    assert np.array_equal(oracle.newell_normal(sq), np.float32([0, 0, -1]))


def _cube_faces():
    normals = np.float32([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]])
    f0 = np.float32([[1, 0, 0], [0, 0, 0], [0, 1, 0], [0, 0, 0], [0, 0, 1], [0, 0, 0]])
    return normals, f0


def test_pt_in_cell_known_answers(oracle):
    n, f = _cube_faces()
    eps = np.float32(1e-4)
    assert oracle.pt_in_cell([.5, .5, .5], n, f, eps) == 1
    assert oracle.pt_in_cell([1.5, .5, .5], n, f, eps) == 0
    assert oracle.pt_in_cell([1 + 5e-5, .5, .5], n, f, eps) == 1      # eps band
    assert oracle.pt_in_cell([1 + 2e-4, .5, .5], n, f, eps) == 0
    assert oracle.pt_in_cell([1.0, 1.0, 1.0], n, f, eps) == 1         # on a corner
    # only significant faces vote: all-positive distances agree -> "inside"
    assert oracle.pt_in_cell([2, 2, 2], n[[0, 2, 4]], f[[0, 2, 4]], eps) == 1


def test_cic_closed_form(oracle):
    rng = np.random.default_rng(3)
    step = np.float32([0.1, 0.1, 0.1])
    gmin = np.float32([0, 0, 0])
    n = 0
    while n < 200:
        pt = (rng.random(3) * 0.8 + 0.05).astype(np.float32)
        fr = pt.astype(np.float64) / 0.1 - np.floor(pt.astype(np.float64) / 0.1)
        if np.any(np.minimum(fr, 1 - fr) * 0.1 < 3e-4):
            continue    # within eps of a grid line: the per-corner nudge applies
        n += 1
        idx, m = oracle.distribute_cic(pt, 2.0, step, gmin, 1e-4)
        base = np.floor(pt.astype(np.float64) / 0.1).astype(int)
        f = pt.astype(np.float64) / step.astype(np.float64) - base
        k = 0
        for z in range(2):
            for y in range(2):
                for x in range(2):
                    assert list(idx[k]) == [base[0] + x, base[1] + y, base[2] + z]
                    w = ((f[0] if x else 1 - f[0]) * (f[1] if y else 1 - f[1]) *
                         (f[2] if z else 1 - f[2]))
                    assert abs(m[k] - 2.0 * w) < 2e-5
                    k += 1
        assert abs(m.astype(np.float64).sum() - 2.0) < 1e-5


def test_grid_step_params_closed_form(oracle):
    gmin, gmax, step = oracle.grid_step_params([0, 0, 0], [1, 1, 1], (64, 64, 64))
    assert np.array_equal(gmin, np.float32([0, 0, 0]))
    assert np.array_equal(gmax, np.float32([1, 1, 1]))
    assert np.array_equal(step, np.float32([1, 1, 1]) / np.float32(63))
    gmin, gmax, step = oracle.grid_step_params([0, 0, 0], [2, 1, 0.5], (5, 5, 5))
    assert np.array_equal(gmin, np.float32([0, -0.5, -0.75]))
    assert np.array_equal(gmax, np.float32([2, 1.5, 1.25]))
    assert np.array_equal(step, np.float32([0.5, 0.5, 0.5]))
    gmin, gmax, step = oracle.grid_step_params([0, 0, 0], [1, 1, 1], (3, 3, 3),
                                               [-1, -1, -1], [3, 3, 3], 2)
    assert np.array_equal(gmin, np.float32([-1, -1, 0]))
    assert np.array_equal(step, np.float32([2, 2, 0.5]))


def test_block_grid_params(oracle):
    g = (65, 65, 65)
    gmin, _, step = oracle.grid_step_params([0, 0, 0], [1, 1, 1], g)
    mn, mx, nm = oracle.block_grid_params([0, 0, 0], [1, 1, 1], gmin, step, 1e-4,
                                          [0, 0, 0], [1, 1, 1], g)
    assert list(mn) == [0, 0, 0] and list(mx) == [64, 64, 64] and list(nm) == [65, 65, 65]
    # two blocks sharing the x = 0.5 face: the lower block gives up the face
    a = oracle.block_grid_params([0, 0, 0], [.5, 1, 1], gmin, step, 1e-4, [0, 0, 0],
                                 [1, 1, 1], g)
    b = oracle.block_grid_params([.5, 0, 0], [1, 1, 1], gmin, step, 1e-4, [0, 0, 0],
                                 [1, 1, 1], g)
    assert a[0][0] == 0 and a[1][0] == 31 and b[0][0] == 32 and b[1][0] == 64


def _p2_block(oracle):
    rng = np.random.default_rng(1234)
    p = rng.random((4096, 3)).astype(np.float32)
    hb = ingest.make_block(p, 4096, (0, 0, 0), (1, 1, 1))
    return hb, oracle.Block.from_host(hb)


def _brute_ptincell(oracle, ob, c, gmin, step, eps):
    bmin, bmax, nrm, f0 = oracle.cell_bounds(ob, c)
    lo = ((bmin - gmin) / step).astype(np.int32)
    hi = ((bmax - gmin) / step).astype(np.int32)
    lop = (lo.astype(np.float32) * step + gmin).astype(np.float32)
    zz, yy, xx = np.meshgrid(np.arange(hi[2] - lo[2] + 1), np.arange(hi[1] - lo[1] + 1),
                             np.arange(hi[0] - lo[0] + 1), indexing="ij")
    px = lop[0] + xx.astype(np.float32) * step[0]
    py = lop[1] + yy.astype(np.float32) * step[1]
    pz = lop[2] + zz.astype(np.float32) * step[2]
    pos = np.zeros(px.shape, bool)
    neg = np.zeros(px.shape, bool)
    for k in range(nrm.shape[0]):
        d = nrm[k, 0] * (px - f0[k, 0]) + nrm[k, 1] * (py - f0[k, 1]) + \
            nrm[k, 2] * (pz - f0[k, 2])
        sig = np.abs(d) > eps
        pos |= sig & (d >= 0)
        neg |= sig & (d < 0)
    ins = ~(pos & neg)
    pts = np.stack([xx[ins] + lo[0], yy[ins] + lo[1], zz[ins] + lo[2]], 1)
    return set(map(tuple, pts.tolist())), lo, hi


def test_reference_scan_counts(oracle):
    """Counts measured from the reference's own scan (built against stand-in
    MPI/DIY headers) on this input: corroboration of the oracle."""
    gold = json.load(open(os.path.join(GOLDEN, "reference_scan_counts.json")))["expected"]
    hb, ob = _p2_block(oracle)
    r = oracle.dense_single(ob, (64, 64, 64))
    st = r["status"]
    gmin, step, eps = r["grid_phys_mins"], r["grid_step_size"], np.float32(1e-4)
    owner = oracle.exact_owner(hb.particles, (64, 64, 64), step, gmin).reshape(64, 64, 64)
    brute = scan = exact = identical = outside = 0
    deposited = np.nonzero(st == oracle.CELL_SCAN)[0]
    assert (st == oracle.CELL_CIC).sum() == 0
    for c in deposited:
        bs, lo, hi = _brute_ptincell(oracle, ob, c, gmin, step, eps)
        _, idx, _ = oracle.cell_points(ob, r["args"], c)
        ss = set(map(tuple, idx.tolist()))
        brute += len(bs)
        scan += len(ss)
        outside += len(ss - bs)
        identical += ss == bs
        exact += int((owner[lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1] == c).sum())
    assert len(deposited) == gold["deposited_cells"]
    assert exact == gold["exact_voronoi_points_of_deposited_cells"]
    assert brute - exact == gold["ptincell_minus_exact_eps_band_points"]
    assert brute - scan == gold["scan_missing_vs_ptincell"]
    assert identical == gold["cells_scan_equals_ptincell"]
    assert outside == gold["scan_points_outside_ptincell"]


def test_mass_conservation_periodic(oracle):
    p = synthetic.uniform(4096, seed=21)
    hb = ingest.periodic_block(p, 4.0 / 16)
    ob = oracle.Block.from_host(hb)
    for g in (16, 48):
        r = oracle.dense_single(ob, (g, g, g), mass=2.0)
        ncell = int(((r["status"] == 0) | (r["status"] == 1)).sum())
        div = np.float64(np.prod(r["grid_step_size"].astype(np.float64)))
        tot = r["density"].astype(np.float64).sum() * div
        assert len(r["offblock"][1]) == 0
        assert abs(tot - 2.0 * ncell) < 1e-4 * 2.0 * ncell
    assert ((r["status"] == 1).sum() == 0)  # fine grid: no CIC fallback


def test_cic_fallback_on_coarse_grid(oracle):
    p = synthetic.uniform(4096, seed=22)
    hb = ingest.periodic_block(p, 4.0 / 16)
    ob = oracle.Block.from_host(hb)
    r = oracle.dense_single(ob, (8, 8, 8))
    assert (r["status"] == oracle.CELL_CIC).sum() > 100
    assert (r["npts"][r["status"] == oracle.CELL_CIC] == 8).all()


def test_exact_owner_brute_force(oracle):
    rng = np.random.default_rng(9)
    p = rng.random((300, 3)).astype(np.float32)
    # planted exact ties: three sites 2^-7 from the grid point (0.5, 0.5, 0.5),
    # nothing else within 0.05 of it
    far = np.linalg.norm(p - 0.5, axis=1) > 0.05
    p[~far] += 0.2
    h = 2.0 ** -7
    p[10] = [0.5, 0.5, 0.5 - h]
    p[200] = [0.5, 0.5, 0.5 + h]
    p[5] = [0.5, 0.5 - h, 0.5]
    p = p.astype(np.float32)
    g = (9, 9, 9)
    step = np.float32([0.125, 0.125, 0.125])
    gmin = np.float32([0, 0, 0])
    own = oracle.exact_owner(p, g, step, gmin)
    zz, yy, xx = np.meshgrid(*(np.arange(9),) * 3, indexing="ij")
    gp = np.stack([xx, yy, zz], -1).reshape(-1, 3).astype(np.float32) * step + gmin
    d = ((gp[:, None, :].astype(np.float64) - p[None].astype(np.float64)) ** 2)
    d2 = (d[..., 0] + d[..., 1]) + d[..., 2]
    brute = np.argmin(d2, axis=1)   # first minimum = lowest id
    assert np.array_equal(own, brute)
    assert own[(4 * 9 + 4) * 9 + 4] == 5


def test_exact_dense_oracle_invariants(oracle):
    """Exact-Voronoi FP64 oracle: every owner is the brute-force nearest site,
    deposited mass is conserved, and each deposited cell with points owns
    exactly npts grid points."""
    p = synthetic.uniform(1500, seed=23)
    hb = ingest.periodic_block(p, 4.0 / 1500 ** (1 / 3))
    ob = oracle.Block.from_host(hb)
    g = 20
    r = oracle.exact_dense(ob, (g, g, g), mass=3.0)
    own = r["owner"]
    step, gmin = r["grid_step_size"], r["grid_phys_mins"]
    zz, yy, xx = np.meshgrid(*(np.arange(g),) * 3, indexing="ij")
    gp = (np.stack([xx, yy, zz], -1).reshape(-1, 3).astype(np.float32) * step + gmin)
    P = hb.particles.astype(np.float64)
    pick = np.random.default_rng(0).choice(gp.shape[0], 800, replace=False)
    d = (gp[pick, None, :].astype(np.float64) - P[None]) ** 2
    brute = np.argmin((d[..., 0] + d[..., 1]) + d[..., 2], axis=1)
    assert np.array_equal(own[pick], brute)
    st, n = r["status"], r["npts"]
    dep = (st == 0) | (st == 1)
    for c in np.nonzero(st == 0)[0][:200]:
        assert (own == c).sum() == n[c]
    div = float(step[0] * step[1] * step[2])          # float product, as the reference
    assert abs(r["density"].sum() * div - 3.0 * dep.sum()) < 1e-9 * 3.0 * dep.sum() + 1e-6


def test_multiblock_model_conserves_mass(oracle):
    """Oracle multi-block model: every deposited cell's mass lands in some
    block grid (own or received) unless its point is held by no block."""
    rng = np.random.default_rng(61)
    w = 0.5
    pts, bid, bounds = [], [], []
    for b in range(8):
        lo = np.array([b % 2, (b // 2) % 2, b // 4]) * w
        pts.append((lo + rng.random((700, 3)) * w).astype(np.float32))
        bid.append(np.full(700, b))
        bounds.append((lo.astype(np.float32), (lo + w).astype(np.float32)))
    blocks = ingest.decomposed_blocks(np.concatenate(pts), np.concatenate(bid), bounds, 0.12)
    obs = [oracle.Block.from_host(b) for b in blocks]
    r = oracle.dense_multi(obs, (33, 33, 33), mass=2.0)
    div = float(r["grid_step_size"][0] * r["grid_step_size"][1] * r["grid_step_size"][2])
    tot = sum(d.astype(np.float64).sum() for d in r["density"]) * div
    ncell = sum(int(((s == 0) | (s == 1)).sum()) for s in r["status"])
    assert r["dropped"] == 0
    assert abs(tot - 2.0 * ncell) < 1e-4 * 2.0 * ncell
    # the block grids partition the global grid
    cover = np.zeros((33, 33, 33), int)
    for mn, mx, nm in r["grids"]:
        cover[mn[2]:mx[2] + 1, mn[1]:mx[1] + 1, mn[0]:mx[0] + 1] += 1
    assert (cover == 1).all()


def test_fp64_density_of_the_scan(oracle):
    """The FP64 accumulation of the scan's point sets: mass conservation in
    closed form (sum rho * h^3 = mass x deposited cells, to double rounding)
    and the float32 density within float rounding of it, zero at the same
    points."""
    p = synthetic.grf_lattice(16, 0.3, seed=8)
    hb = ingest.periodic_block(p, 4.0 / 16)
    ob = oracle.Block.from_host(hb)
    g = (32, 32, 32)
    r = oracle.dense_single(ob, g, mass=1.5)
    d64 = oracle.dense_single_fp64(ob, g, mass=1.5)
    ncell = int(((r["status"] == 0) | (r["status"] == 1)).sum())
    h3 = np.prod(r["grid_step_size"].astype(np.float64))
    assert len(r["offblock"][1]) == 0
    assert abs(d64.sum() * h3 - 1.5 * ncell) <= 1e-12 * 1.5 * ncell
    e = oracle.rel_error(r["density"], d64)
    assert e["zeros_identical"] and e["points"] > 0
    # each point is a chain of float(double + double) roundings: a few ulp
    assert 0 < e["max_rel"] < 1e-6 and e["rms_rel"] < 1e-7
