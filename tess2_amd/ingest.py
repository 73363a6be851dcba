"""Host-side ingestion: particles -> Qhull Delaunay -> tet_t arrays.

The reference computes each block's Delaunay tessellation with Qhull
("qhull d Qt", src/tess-qhull.c:24-73), converts facets to tet_t with
tets[i] opposite verts[i] (src/tess-qhull.c:83-225) and picks, for every
vertex, the last tet that contains it (fill_vert_to_tet, src/tess.cpp:767-787).

Here Qhull is SciPy's bundled copy, called as Delaunay(points, "Qt"): SciPy
prefixes "d" and documents neighbors[j] as opposite simplices[j], the tet_t
convention.  Periodic images ("ghosts") are appended after the original
particles, as the reference's ghost exchange appends received particles
(include/tess/delaunay.h:44-49, wrap_pt src/tess.cpp:703-710).

For large inputs the tessellation is computed per x-y column in worker
processes, each column with a margin of neighbouring particles; a tet is
kept by the column holding its lowest-id vertex, only if its circumsphere
(clipped to the cloud's bounding box) lies inside that column's padded
region, so that it is Delaunay for the whole set; neighbour links are then
rebuilt across columns by matching shared faces.  Generic (non-cospherical)
inputs have a unique Delaunay tessellation, so the merge equals a single
Qhull call up to tet numbering.
"""
from __future__ import annotations

import os
import sys
import time
from dataclasses import dataclass, field

import numpy as np


@dataclass
class HostBlock:
    """One block: the dblock_t / DBlock fields the density path reads."""
    particles: np.ndarray          # (num_particles, 3) float32, originals first
    tets: np.ndarray               # (num_tets, 8) int32: verts[4], tets[4]
    vert_to_tet: np.ndarray        # (num_particles,) int32
    num_orig: int
    bounds_min: np.ndarray         # (3,) float32
    bounds_max: np.ndarray         # (3,) float32
    gid: int = 0
    ghost_src: np.ndarray | None = None   # original id of each ghost image
    ghost_gid: np.ndarray | None = None   # gid of the block holding each ghost as an original
    ghost_cell: np.ndarray | None = None  # and its index among that block's originals
    info: dict = field(default_factory=dict)

    @property
    def num_particles(self):
        return self.particles.shape[0]

    @property
    def num_tets(self):
        return self.tets.shape[0]


def vert_to_tet(tets: np.ndarray, num_particles: int) -> np.ndarray:
    """For each vertex the LAST tet (highest index) containing it, -1 if none
    (src/tess.cpp:774-786: the last one wins)."""
    v2t = np.full(num_particles, -1, np.int32)
    T = tets.shape[0]
    if T == 0:
        return v2t
    verts = tets[:, :4]
    ids = np.broadcast_to(np.arange(T, dtype=np.int32)[:, None], verts.shape)
    # the last tet containing a vertex is the one with the largest index
    np.maximum.at(v2t, verts.ravel(), ids.ravel())
    return v2t


def qhull_tets(points: np.ndarray, qhull_points: np.ndarray | None = None) -> np.ndarray:
    """Delaunay tets of float32 points as an (T, 8) int32 tet_t array.
    Points are promoted to double as the reference does (src/tess-qhull.c:39-42).
    qhull_points: float64 positions of the same particles that decide the
    triangulation instead (for float32 sets too degenerate for Qhull)."""
    from scipy.spatial import Delaunay
    if qhull_points is not None:
        p64 = np.ascontiguousarray(qhull_points, dtype=np.float64)
    else:
        p64 = np.ascontiguousarray(points, dtype=np.float32).astype(np.float64)
    d = Delaunay(p64, qhull_options="Qt")
    T = d.simplices.shape[0]
    tets = np.empty((T, 8), np.int32)
    tets[:, :4] = d.simplices
    tets[:, 4:] = d.neighbors
    return tets


def periodic_ghosts(p: np.ndarray, width: float, lo=0.0, hi=1.0):
    """Periodic images of p within `width` of the box [lo, hi)^3, shifted in
    float32 as wrap_pt does (src/tess.cpp:703-710).  Returns (ghosts, src)."""
    p = np.ascontiguousarray(p, dtype=np.float32)
    ext = np.float32(hi - lo)
    w_lo = np.float32(lo - width)
    w_hi = np.float32(hi + width)
    gs, src = [], []
    ids = np.arange(p.shape[0], dtype=np.int64)
    # candidates: only particles within width of some face can have images
    near = np.any((p < np.float32(lo + width)) | (p > np.float32(hi - width)), axis=1)
    pn, idn = p[near], ids[near]
    for sx in (-1, 0, 1):
        for sy in (-1, 0, 1):
            for sz in (-1, 0, 1):
                if sx == sy == sz == 0:
                    continue
                s = np.array([sx, sy, sz], np.float32) * ext
                q = (pn + s).astype(np.float32)
                m = np.all((q >= w_lo) & (q <= w_hi), axis=1)
                gs.append(q[m])
                src.append(idn[m])
    if gs:
        return np.concatenate(gs).astype(np.float32), np.concatenate(src)
    return np.zeros((0, 3), np.float32), np.zeros(0, np.int64)


def _circumspheres(pts64: np.ndarray, tv: np.ndarray):
    """Circumcenters and squared radii in float64 (for the slab acceptance test)."""
    a = pts64[tv[:, 0]]
    b = pts64[tv[:, 1]] - a
    c = pts64[tv[:, 2]] - a
    d = pts64[tv[:, 3]] - a
    bb = (b * b).sum(1)
    cc = (c * c).sum(1)
    dd = (d * d).sum(1)
    cxd = np.cross(c, d)
    dxb = np.cross(d, b)
    bxc = np.cross(b, c)
    den = 2.0 * (b * cxd).sum(1)
    with np.errstate(divide="ignore", invalid="ignore"):
        off = (bb[:, None] * cxd + cc[:, None] * dxb + dd[:, None] * bxc) / den[:, None]
    r2 = (off * off).sum(1)
    return a + off, r2


def _column_job(args):
    """Delaunay of one x-y column plus its margin; keeps the tets the column
    owns (lowest-id vertex inside it) whose circumsphere, clipped to the
    point cloud's bounding box, lies inside the padded column: no point
    outside the padded column can then lie in the sphere, so the tet is
    Delaunay for the whole set."""
    (pts, lo, hi, margin, cloud_lo, cloud_hi, q64) = args
    if q64 is not None:
        pts = q64          # the float64 positions decide membership too
    pad_lo = lo - margin
    pad_hi = hi + margin
    inside = np.ones(pts.shape[0], bool)
    for ax in (0, 1):
        inside &= (pts[:, ax] >= pad_lo[ax]) & (pts[:, ax] < pad_hi[ax])
    sel = np.nonzero(inside)[0]
    sub = pts[sel]
    t0 = time.time()
    tets = qhull_tets(sub, None if q64 is None else sub)
    tv = sel[tets[:, :4]]                        # global vertex ids
    ov = pts[tv.min(1)]
    own = np.ones(tv.shape[0], bool)
    for ax in (0, 1):
        own &= (ov[:, ax] >= lo[ax]) & (ov[:, ax] < hi[ax])
    tv = tv[own]
    cen, r2 = _circumspheres(pts.astype(np.float64), tv)
    # flat tets (cospherical points) have no finite sphere and are never kept
    with np.errstate(invalid="ignore", over="ignore"):
        r = np.sqrt(r2)
        # distance of the centre to the cloud box along each axis
        gap = np.maximum(np.maximum(cloud_lo - cen, cen - cloud_hi), 0.0)
        ok = np.isfinite(r)
        for ax in (0, 1):
            # half-width along ax of the sphere's intersection with the box
            rest = r2 - sum(gap[:, o] ** 2 for o in range(3) if o != ax)
            hw = np.sqrt(np.maximum(rest, 0.0))
            smin = np.maximum(cen[:, ax] - hw, cloud_lo[ax])
            smax = np.minimum(cen[:, ax] + hw, cloud_hi[ax])
            ok &= (smin >= pad_lo[ax]) & (smax <= pad_hi[ax])
    return tv, ok, time.time() - t0, sub.shape[0]


def _rebuild_neighbors(tv: np.ndarray) -> np.ndarray:
    """tets[t][i] = tet sharing the face opposite verts[i] (-1 if none)."""
    T = tv.shape[0]
    faces = np.empty((T, 4, 3), np.int64)
    for i in range(4):
        faces[:, i] = np.delete(tv, i, axis=1)
    faces.sort(axis=2)
    f = faces.reshape(-1, 3)
    n = np.int64(int(tv.max()) + 1)
    key = (f[:, 0] * n + f[:, 1]) * n + f[:, 2]
    order = np.argsort(key, kind="stable")
    ks = key[order]
    nb = np.full(T * 4, -1, np.int64)
    same = ks[1:] == ks[:-1]
    i0 = order[:-1][same]
    i1 = order[1:][same]
    if np.any(ks[2:][same[1:] & same[:-1]] == ks[1:-1][same[1:] & same[:-1]]):
        raise RuntimeError("face shared by more than two tets: inconsistent slabs")
    nb[i0] = i1 // 4
    nb[i1] = i0 // 4
    return nb.reshape(T, 4).astype(np.int32)


def open_star_vertices(tets: np.ndarray, num_particles: int) -> np.ndarray:
    """Mask of vertices on a boundary face (a face with no neighbour): exactly
    the vertices whose star is not closed, i.e. complete() would fail."""
    m = np.zeros(num_particles, bool)
    for i in range(4):
        b = tets[:, 4 + i] < 0
        for j in range(4):
            if j != i:
                m[tets[b, j]] = True
    return m


def tessellate(points: np.ndarray, workers: int = 1, columns: int | None = None,
               log=None, min_points: int = 200_000, closed_first: int = 0,
               qhull_points: np.ndarray | None = None) -> np.ndarray:
    """Delaunay tets (T, 8) of all points.  With workers > 1 the points are
    split into x-y columns tessellated in parallel processes and merged into
    one tessellation (tet numbering then follows the columns).  Only tets
    certified Delaunay for the whole set are kept; the first `closed_first`
    points (the original particles of a ghosted block) must end up with
    closed stars, otherwise the margin doubles and the columns are redone.
    Degenerate input whose columns cannot be merged falls back to one call.
    qhull_points: float64 positions that decide the triangulation (see
    qhull_tets); the columns are then cut and certified on them as well."""
    points = np.ascontiguousarray(points, dtype=np.float32)
    q64 = None if qhull_points is None else np.ascontiguousarray(qhull_points, np.float64)
    n = points.shape[0]
    if workers <= 1 or n < min_points:
        return qhull_tets(points, q64)
    from concurrent.futures import ProcessPoolExecutor
    cols = columns or workers
    nx = int(np.floor(np.sqrt(cols)))
    while cols % nx:
        nx -= 1
    ny = cols // nx
    geo = points if q64 is None else q64
    qx = np.quantile(geo[:, 0], np.linspace(0, 1, nx + 1))
    qy = np.quantile(geo[:, 1], np.linspace(0, 1, ny + 1))
    qx[0], qx[-1], qy[0], qy[-1] = -np.inf, np.inf, -np.inf, np.inf
    cloud_lo = geo.min(0).astype(np.float64)
    cloud_hi = geo.max(0).astype(np.float64)
    extent = float((cloud_hi - cloud_lo).max())
    spacing = extent / max(n, 1) ** (1.0 / 3.0)
    margin = 3.0 * spacing
    for attempt in range(5):
        jobs = [(points, np.array([qx[i], qy[j]]), np.array([qx[i + 1], qy[j + 1]]),
                 margin, cloud_lo, cloud_hi, q64) for i in range(nx) for j in range(ny)]
        parts, dropped = [], 0
        with ProcessPoolExecutor(max_workers=workers) as ex:
            for k, (tv, ok, dt, m) in enumerate(ex.map(_column_job, jobs)):
                if log:
                    log(f"  column {k + 1}/{len(jobs)}: {m} points, {int(ok.sum())} tets, "
                        f"{dt:.1f}s")
                dropped += int((~ok).sum())
                parts.append(tv[ok])
        tv = np.concatenate(parts).astype(np.int32)
        try:
            nb = _rebuild_neighbors(tv)
        except RuntimeError:
            # cospherical point sets have no unique Delaunay triangulation:
            # columns can split the same empty sphere differently, so the
            # pieces do not fit; one Qhull call decides consistently
            if log:
                log("  columns disagree on degenerate (cospherical) cells; one Qhull call")
            return qhull_tets(points, q64)
        tets = np.empty((tv.shape[0], 8), np.int32)
        tets[:, :4] = tv
        tets[:, 4:] = nb
        if closed_first <= 0 and dropped == 0:
            return tets
        if closed_first > 0:
            bad = int(open_star_vertices(tets, n)[:closed_first].sum())
            if bad == 0:
                if log:
                    log(f"  merged {tets.shape[0]} tets ({dropped} uncertified hull-side "
                        f"tets left out); all {closed_first} original stars closed")
                return tets
            if log:
                log(f"  {bad} original stars open; margin x2")
        margin *= 2.0
    raise RuntimeError("column tessellation did not converge")


def make_block(particles: np.ndarray, num_orig: int, bounds_min, bounds_max,
               workers: int = 1, gid: int = 0, ghost_src=None, log=None,
               closed_first: int = 0, qhull_points=None) -> HostBlock:
    particles = np.ascontiguousarray(particles, dtype=np.float32)
    t0 = time.time()
    tets = tessellate(particles, workers=workers, log=log, closed_first=closed_first,
                      qhull_points=qhull_points)
    t1 = time.time()
    v2t = vert_to_tet(tets, particles.shape[0])
    return HostBlock(particles=particles, tets=tets, vert_to_tet=v2t, num_orig=num_orig,
                     bounds_min=np.asarray(bounds_min, np.float32),
                     bounds_max=np.asarray(bounds_max, np.float32), gid=gid,
                     ghost_src=ghost_src, info=dict(qhull_s=t1 - t0))


def ghost_positions64(p: np.ndarray, p64: np.ndarray, g: np.ndarray, src: np.ndarray):
    """float64 positions of the periodic images g of p (sources src), from
    the float64 positions p64 of the originals: the same whole-box shift."""
    shift = np.rint(g.astype(np.float64) - p[src].astype(np.float64))
    return np.concatenate([p64, p64[src] + shift])


def periodic_block(p: np.ndarray, ghost_width: float, workers: int = 1,
                   log=None, p64: np.ndarray | None = None) -> HostBlock:
    """One periodic unit-cube block: originals + periodic ghost images.
    p64: float64 positions of p that decide the triangulation (qhull_tets)."""
    p = np.ascontiguousarray(p, dtype=np.float32)
    g, src = periodic_ghosts(p, ghost_width)
    allp = np.concatenate([p, g]).astype(np.float32)
    q64 = None if p64 is None else ghost_positions64(p, p64, g, src)
    return make_block(allp, p.shape[0], (0, 0, 0), (1, 1, 1), workers=workers,
                      ghost_src=src, log=log, closed_first=p.shape[0], qhull_points=q64)


def default_workers() -> int:
    n = os.cpu_count() or 1
    return max(1, min(16, n))


def _block_job(args):
    allp, norig, lo, hi, gid = args
    t0 = time.time()
    tets = qhull_tets(allp)
    return tets, time.time() - t0


def decomposed_blocks(points: np.ndarray, block_id: np.ndarray, bounds, ghost_width: float,
                      workers: int = 1, log=None) -> list:
    """DIY-style blocks: block b holds the particles with block_id == b (its
    original particles, in input order) followed by the particles of the
    other blocks within ghost_width of its bounds (its ghosts), and its own
    Qhull tessellation (src/tess.cpp:351-457 computes one per block after the
    ghost exchange).  Blocks are tessellated in parallel processes.
    bounds: list of (lo, hi) per block; gid = block index."""
    points = np.ascontiguousarray(points, dtype=np.float32)
    block_id = np.asarray(block_id)
    w = np.float32(ghost_width)
    # index of each particle among its own block's originals
    home_idx = np.zeros(points.shape[0], np.int32)
    for b in range(len(bounds)):
        sel = block_id == b
        home_idx[sel] = np.arange(int(sel.sum()), dtype=np.int32)
    jobs, homes = [], []
    for b, (lo, hi) in enumerate(bounds):
        lo = np.asarray(lo, np.float32)
        hi = np.asarray(hi, np.float32)
        own = points[block_id == b]
        near = (block_id != b) & np.all(points >= lo - w, axis=1) & np.all(points <= hi + w, axis=1)
        allp = np.concatenate([own, points[near]]).astype(np.float32)
        jobs.append((allp, own.shape[0], lo, hi, b))
        homes.append((block_id[near].astype(np.int32), home_idx[near]))
    if workers > 1 and len(jobs) > 1:
        from concurrent.futures import ProcessPoolExecutor
        with ProcessPoolExecutor(max_workers=workers) as ex:
            results = list(ex.map(_block_job, jobs))
    else:
        results = [_block_job(j) for j in jobs]
    blocks = []
    for (allp, norig, lo, hi, b), (tets, dt), (hg, hc) in zip(jobs, results, homes):
        v2t = vert_to_tet(tets, allp.shape[0])
        blocks.append(HostBlock(particles=allp, tets=tets, vert_to_tet=v2t, num_orig=norig,
                                bounds_min=lo, bounds_max=hi, gid=b, ghost_gid=hg,
                                ghost_cell=hc, info=dict(qhull_s=dt)))
        if log:
            log(f"  block {b}: {norig} + {allp.shape[0] - norig} ghosts, {tets.shape[0]} tets, "
                f"{dt:.1f}s")
    return blocks


def adjacent_blocks(bounds, b) -> list:
    """Blocks whose closed boxes touch block b's (face, edge or corner
    neighbours of a regular decomposition), ascending index."""
    lo, hi = (np.asarray(x, np.float32) for x in bounds[b])
    out = []
    for n, (l2, h2) in enumerate(bounds):
        if n != b and np.all(np.asarray(l2) <= hi) and np.all(np.asarray(h2) >= lo):
            out.append(n)
    return out


def exchanged_blocks(points: np.ndarray, block_id: np.ndarray, bounds, select,
                     workers: int = 1, log=None, max_rounds: int = 16):
    """DIY-style blocks whose ghosts come from the reference's exchange
    rounds, non-periodic (tess(), src/tess.cpp:45-116; delaunay() :351-457):
    every block starts from its own particles; each round tessellates every
    block (originals + the ghosts received so far) and asks `select` which
    originals must go to which neighbour -- the test of incomplete_cells()
    (:492-628) -- then sends those a neighbour does not hold yet.  A block's
    links start as its adjacent blocks; after each round they grow by the
    neighbours' own adjacent blocks, and the next round tests only the new
    ones (the reference's last_neighbor).  Stops after a round that sends
    nothing.  select(block, boxes, first) -> uint64 mask per original
    particle (bit i: neighbour i of boxes, an n x 6 array of min / max).
    Returns (blocks with ghost homes, rounds, particles sent per round)."""
    points = np.ascontiguousarray(points, dtype=np.float32)
    block_id = np.asarray(block_id)
    nb = len(bounds)
    own = [np.nonzero(block_id == b)[0] for b in range(nb)]
    home_idx = np.zeros(points.shape[0], np.int32)
    for b in range(nb):
        home_idx[own[b]] = np.arange(own[b].size, dtype=np.int32)
    adj = [adjacent_blocks(bounds, b) for b in range(nb)]
    links = [list(a) for a in adj]
    last = [0] * nb
    ghosts = [[] for _ in range(nb)]
    have = [set() for _ in range(nb)]
    box = np.array([[*np.asarray(lo, np.float32), *np.asarray(hi, np.float32)]
                    for lo, hi in bounds], np.float32)
    sent_per_round = []
    blocks = []
    for rnd in range(1, max_rounds + 1):
        jobs = []
        for b in range(nb):
            idx = np.concatenate([own[b], np.asarray(ghosts[b], np.int64)]).astype(np.int64)
            lo, hi = (np.asarray(x, np.float32) for x in bounds[b])
            jobs.append((points[idx], own[b].size, lo, hi, b))
        if workers > 1 and nb > 1:
            from concurrent.futures import ProcessPoolExecutor
            with ProcessPoolExecutor(max_workers=workers) as ex:
                results = list(ex.map(_block_job, jobs))
        else:
            results = [_block_job(j) for j in jobs]
        blocks = []
        for (allp, norig, lo, hi, b), (tets, dt) in zip(jobs, results):
            g = np.asarray(ghosts[b], np.int64)
            blocks.append(HostBlock(particles=allp, tets=tets,
                                    vert_to_tet=vert_to_tet(tets, allp.shape[0]), num_orig=norig,
                                    bounds_min=lo, bounds_max=hi, gid=b,
                                    ghost_gid=block_id[g].astype(np.int32),
                                    ghost_cell=home_idx[g], info=dict(qhull_s=dt, round=rnd)))
        sent = 0
        new = [[] for _ in range(nb)]
        for b in range(nb):
            if last[b] >= len(links[b]):
                continue
            masks = np.asarray(select(blocks[b], box[links[b]], last[b]), np.uint64)
            for p in np.nonzero(masks)[0]:
                m = int(masks[p])
                gidx = int(own[b][p])
                for i in range(last[b], len(links[b])):
                    if (m >> i) & 1:
                        d = links[b][i]
                        if gidx not in have[d]:
                            have[d].add(gidx)
                            new[d].append(gidx)
                            sent += 1
        sent_per_round.append(sent)
        if log:
            log(f"  round {rnd}: {sent} particles sent")
        if sent == 0:
            return blocks, rnd, sent_per_round
        for d in range(nb):
            ghosts[d].extend(new[d])
        # links grow by the neighbours' adjacent blocks; only those are tested next
        for b in range(nb):
            grown = sorted(set(n2 for n in links[b] for n2 in adj[n]) - set(links[b]) - {b})
            last[b] = len(links[b])
            links[b].extend(grown)
            if len(links[b]) > 64:
                raise ValueError("more than 64 neighbour links")
    raise RuntimeError(f"ghost exchange did not settle in {max_rounds} rounds")


def _sorted_keys(t: np.ndarray):
    """Rows of sorted vertex ids as two int64 keys each, sorted, unique."""
    if t.shape[0] == 0:
        return np.zeros((0, 2), np.int64)
    n = np.int64(t.max()) + 1
    k = np.stack([t[:, 0] * n + t[:, 1], t[:, 2] * n + t[:, 3]], axis=1)
    o = np.lexsort((k[:, 1], k[:, 0]))
    k = k[o]
    keep = np.ones(k.shape[0], bool)
    keep[1:] = np.any(k[1:] != k[:-1], axis=1)
    return k[keep]


def stars_closed(blocks, points, block_id, workers: int = 1, tets=None) -> dict:
    """Whether every original particle's star in its block equals its star in
    the Delaunay tessellation of all points (tets, computed when not given):
    the tets with at least one of a block's originals, as sorted global
    vertex ids, are compared as sets."""
    points = np.ascontiguousarray(points, dtype=np.float32)
    block_id = np.asarray(block_id)
    if tets is None:
        tets = tessellate(points, workers=workers)
    gt = np.sort(tets[:, :4].astype(np.int64), axis=1)
    bad_blocks, checked = [], 0
    for hb in blocks:
        b = hb.gid
        own = np.nonzero(block_id == b)[0]
        gmap = np.concatenate([own, _ghost_globals(hb, block_id)])
        lt = np.sort(gmap[hb.tets[:, :4]], axis=1)
        mine = np.zeros(points.shape[0], bool)
        mine[own] = True
        want = _sorted_keys(gt[mine[gt].any(axis=1)])
        got = _sorted_keys(lt[mine[lt].any(axis=1)])
        checked += want.shape[0]
        if want.shape != got.shape or not np.array_equal(want, got):
            bad_blocks.append(b)
    return {"closed": not bad_blocks, "bad_blocks": bad_blocks, "tets_checked": int(checked)}


def _ghost_globals(hb, block_id):
    """Global indices of a block's ghosts from their homes (gid, index)."""
    block_id = np.asarray(block_id)
    if hb.ghost_gid is None or len(hb.ghost_gid) == 0:
        return np.zeros(0, np.int64)
    out = np.zeros(len(hb.ghost_gid), np.int64)
    for g in np.unique(hb.ghost_gid):
        idx = np.nonzero(block_id == g)[0]
        sel = hb.ghost_gid == g
        out[sel] = idx[hb.ghost_cell[sel]]
    return out


def stars_certified(blocks, points, block_id, workers: int = 1, log=None) -> dict:
    """A certificate that every original particle's star in its block is its
    Delaunay star among all points, without tessellating all points: every
    tet of the star has an empty circumsphere (its 5 nearest points to the
    circumcenter, in float64, include no point strictly inside beyond a 1e-9
    relative margin), and every face of the star on the block tessellation's
    hull is a facet of the convex hull of all points (so the star is closed
    wherever the global one is).  For points in general position such a star
    is the Delaunay star."""
    from scipy.spatial import ConvexHull, cKDTree
    points = np.ascontiguousarray(points, dtype=np.float32)
    block_id = np.asarray(block_id)
    p64 = points.astype(np.float64)
    tree = cKDTree(p64)
    hull = {tuple(f) for f in np.sort(ConvexHull(p64).simplices, axis=1).tolist()}
    bad_blocks, open_faces, nonempty, checked = [], 0, 0, 0
    for hb in blocks:
        b = hb.gid
        own = np.nonzero(block_id == b)[0]
        gmap = np.concatenate([own, _ghost_globals(hb, block_id)])
        sel = (hb.tets[:, :4] < hb.num_orig).any(axis=1)
        tets = hb.tets[sel]
        tv = tets[:, :4]
        p = hb.particles.astype(np.float64)
        a, bb, c, d = (p[tv[:, k]] for k in range(4))
        # circumcenter of each tet in float64: solve 2 (x - d) . [a-d; b-d; c-d] = |.|^2
        m = np.stack([a - d, bb - d, c - d], axis=1)
        rhs = 0.5 * np.einsum("ijk,ijk->ij", m, m)
        ok = np.abs(np.linalg.det(m)) > 0
        cen = np.full((tv.shape[0], 3), np.nan)
        cen[ok] = np.linalg.solve(m[ok], rhs[ok][:, :, None])[:, :, 0] + d[ok]
        rad = np.linalg.norm(cen - d, axis=1)
        dist, idx = tree.query(np.where(ok[:, None], cen, 0.0), k=5, workers=workers)
        gv = gmap[tv]
        inside = (dist < rad[:, None] * (1.0 - 1e-9)) & ~np.any(idx[:, :, None] == gv[:, None, :],
                                                                axis=2)
        n_bad = int(inside[ok].any(axis=1).sum())
        # hull faces of the block tessellation that touch an original
        faces = []
        for j in range(4):
            rows = tets[:, 4 + j] == -1
            if rows.any():
                f = np.delete(tv[rows], j, axis=1)
                f = f[(f < hb.num_orig).any(axis=1)]
                faces.append(np.sort(gmap[f], axis=1))
        n_open = 0
        if faces:
            fl = np.unique(np.concatenate(faces), axis=0)
            n_open = sum(1 for f in fl.tolist() if tuple(f) not in hull)
        open_faces += n_open
        nonempty += n_bad
        checked += int(tv.shape[0])
        if n_open or n_bad:
            bad_blocks.append(b)
        if log:
            log(f"  block {b}: {tv.shape[0]} tets checked, {n_bad} non-empty spheres, "
                f"{n_open} hull faces not on the convex hull")
    return {"closed": not bad_blocks, "bad_blocks": bad_blocks, "nonempty_spheres": nonempty,
            "open_faces_off_hull": open_faces, "tets_checked": checked}
