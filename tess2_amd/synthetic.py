"""Seeded synthetic particle sets for the benchmark and tests.

All coordinates are drawn in float64 with NumPy's PCG64 and rounded to
float32.  Periodic sets live in the unit cube [0, 1)^3.

  uniform(n, seed)                 uniform random particles
  grf_lattice(n_side, rms, seed)   lattice displaced by a Gaussian random field
                                   with P(k) ~ k^-2 (Zel'dovich displacement),
                                   rms displacement `rms` lattice spacings
  jittered_lattice(n_side, jit)    exact lattice + uniform jitter of +-jit spacing
  plummer_blocks(...)              uneven DIY-style blocks of Plummer clusters
"""
from __future__ import annotations

import numpy as np


def _wrap01(x64: np.ndarray) -> np.ndarray:
    x = np.mod(x64, 1.0).astype(np.float32)
    x[x >= np.float32(1.0)] = np.float32(0.0)
    return x


def uniform(n: int, seed: int = 0) -> np.ndarray:
    rng = np.random.Generator(np.random.PCG64(seed))
    return _wrap01(rng.random((n, 3)))


def grf_lattice(n_side: int, rms: float, seed: int = 1, slope: float = -2.0) -> np.ndarray:
    """n_side^3 particles: lattice sites (i+1/2)/n_side displaced by the
    Zel'dovich displacement psi(k) = i k/k^2 delta(k) of a Gaussian random
    field with P(k) ~ k^slope, scaled to an rms |psi| of rms/n_side."""
    rng = np.random.Generator(np.random.PCG64(seed))
    n = n_side
    white = rng.standard_normal((n, n, n))
    wk = np.fft.rfftn(white)
    kx = np.fft.fftfreq(n) * n
    kz = np.fft.rfftfreq(n) * n
    KX, KY, KZ = np.meshgrid(kx, kx, kz, indexing="ij")
    k2 = KX * KX + KY * KY + KZ * KZ
    k2[0, 0, 0] = 1.0
    amp = k2 ** (slope / 4.0)          # sqrt(P(k)) with P ~ k^slope
    amp[0, 0, 0] = 0.0
    dk = wk * amp
    psi = []
    for K in (KX, KY, KZ):
        comp = np.fft.irfftn(1j * K / k2 * dk, s=(n, n, n), axes=(0, 1, 2))
        psi.append(comp)
    psi = np.stack(psi, axis=-1).reshape(-1, 3)
    cur = np.sqrt(np.mean(np.sum(psi * psi, axis=1)))
    psi *= (rms / n) / cur
    g = (np.arange(n) + 0.5) / n
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    lat = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    return _wrap01(lat + psi)


def jittered_lattice64(n_side: int, jitter: float = 1e-6, seed: int = 4) -> np.ndarray:
    """The jittered lattice before rounding to float32, wrapped into [0, 1)
    in float64.  Below float32 resolution the jitter survives only here: a
    tessellation of these positions is one Delaunay triangulation of the
    (degenerate) float32 particles, decided consistently everywhere."""
    rng = np.random.Generator(np.random.PCG64(seed))
    n = n_side
    g = (np.arange(n) + 0.5) / n
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    lat = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    lat += rng.uniform(-jitter, jitter, lat.shape) / n
    return np.mod(lat, 1.0)


def jittered_lattice(n_side: int, jitter: float = 1e-6, seed: int = 4) -> np.ndarray:
    return _wrap01(jittered_lattice64(n_side, jitter, seed))


def lattice_tiebreak(p64: np.ndarray, n_side: int, amp: float = 1e4) -> np.ndarray:
    """Positions that decide the Delaunay triangulation of a jittered
    lattice: each site's jitter scaled by amp (1e-6 -> 1e-2 spacing).  At
    1e-6 spacing Qhull's own precision handling merges the near-cospherical
    cubes and triangulates them arbitrarily, so neighbouring columns cannot
    agree (at 1e-3 spacing still, for 128^3 sites); at 1e-2 spacing the
    input is generic, the triangulation unique,
    and for the float32 particles it is a Delaunay triangulation wherever
    they are exactly cospherical (~75 % of the sites at 128^3 sit exactly on
    the lattice) and differs from one elsewhere only where float32 rounding
    reversed a jitter's sign."""
    lat = (np.round(p64 * n_side - 0.5) + 0.5) / n_side
    return lat + (p64 - lat) * amp


def plummer_blocks(nblocks_side: int = 4, nmin: int = 5000, nmax: int = 200000,
                   scale: float = 0.1, background: float = 0.1, seed: int = 2):
    """nblocks_side^3 blocks tiling [0,1)^3; each holds a log-uniform count of
    particles: a Plummer sphere (scale radius scale*block width, centred in
    the block, truncated to it) plus a `background` fraction uniform in the
    block.  Returns (points, block_id, counts, block_bounds)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    nb = nblocks_side ** 3
    w = 1.0 / nblocks_side
    counts = np.exp(rng.uniform(np.log(nmin), np.log(nmax), nb)).astype(np.int64)
    pts, bid, bounds = [], [], []
    for b in range(nb):
        i, j, k = b % nblocks_side, (b // nblocks_side) % nblocks_side, b // nblocks_side ** 2
        lo = np.array([i, j, k], np.float64) * w
        hi = lo + w
        n = int(counts[b])
        nbg = int(round(background * n))
        ncl = n - nbg
        out = []
        a = scale * w
        while sum(len(o) for o in out) < ncl:
            m = 2 * (ncl - sum(len(o) for o in out)) + 16
            u = rng.random(m)
            r = a / np.sqrt(u ** (-2.0 / 3.0) - 1.0)
            d = rng.standard_normal((m, 3))
            d /= np.linalg.norm(d, axis=1, keepdims=True)
            q = (lo + hi) / 2 + d * r[:, None]
            ok = np.all((q >= lo) & (q < hi), axis=1)
            out.append(q[ok])
        cl = np.concatenate(out)[:ncl]
        bg = lo + rng.random((nbg, 3)) * w
        allp = np.concatenate([cl, bg]).astype(np.float32)
        allp = np.minimum(np.maximum(allp, lo.astype(np.float32)),
                          np.nextafter(hi.astype(np.float32), np.float32(0)))
        pts.append(allp)
        bid.append(np.full(n, b, np.int32))
        bounds.append((lo.astype(np.float32), hi.astype(np.float32)))
    return np.concatenate(pts), np.concatenate(bid), counts, bounds


CONFIGS = {
    0: dict(name="uniform32k_64", n_side=32, grid=64, kind="uniform", ghost=4.0),
    1: dict(name="grf2M_256", n_side=128, grid=256, kind="grf", rms=0.3, ghost=4.0),
    # 4^3 blocks, each on a 64^3 block grid (global 256^3); ghost 0.1 block width
    2: dict(name="plummer64_256", kind="plummer", nblocks_side=4, nmin=5000, nmax=200000,
            grid=256, ghost_blocks=0.1),
    3: dict(name="grf16M_512", n_side=256, grid=512, kind="grf", rms=0.5, ghost=4.0),
    4: dict(name="lattice2M_512", n_side=128, grid=512, kind="lattice", jitter=1e-6,
            ghost=3.0),
}


def config_particles(idx: int, n_side: int | None = None) -> np.ndarray:
    c = CONFIGS[idx]
    ns = n_side or c["n_side"]
    if c["kind"] == "uniform":
        return uniform(ns ** 3, seed=idx)
    if c["kind"] == "grf":
        return grf_lattice(ns, c["rms"], seed=idx)
    if c["kind"] == "lattice":
        return jittered_lattice(ns, c["jitter"], seed=idx)
    if c["kind"] == "plummer":
        raise ValueError("the plummer workload is a block decomposition: use plummer_blocks")
    raise ValueError(idx)
