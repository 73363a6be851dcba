"""Benchmark of the post-tessellation path on one MI355X.

Workload (default): 2,097,152 particles (128^3 lattice displaced by a
Gaussian random field, P(k) ~ k^-2, rms 0.3 spacing, periodic unit cube,
periodic ghost images 4 spacings wide), Qhull Delaunay on the host, then on
the GPU per step: circumcenters -> stars -> volumes -> face planes ->
scanline assignment -> deposit onto a 256^3 float32 grid, all with inputs
resident in HBM.  One JSON line on stdout.

  python bench.py [--steps K] [--warmup W] [--config {0,1}] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("grid pts/s + HBM roofline %, 2.1M particles→256³; "
          "mixed vs FP64 GPU path")
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def setup(cfg, load=None):
    from tess2_amd import ingest, synthetic
    c = synthetic.CONFIGS[cfg]
    if load:
        z = np.load(load)
        hb = ingest.HostBlock(particles=z["particles"], tets=z["tets"],
                              vert_to_tet=z["vert_to_tet"], num_orig=int(z["num_orig"]),
                              bounds_min=z["bounds_min"], bounds_max=z["bounds_max"])
        log(f"[setup] loaded {load}: {hb.num_orig} particles, {hb.num_tets} tets")
        return c, hb
    t0 = time.time()
    p = synthetic.config_particles(cfg)
    g, src = ingest.periodic_ghosts(p, c["ghost"] / c["n_side"])
    allp = np.concatenate([p, g]).astype(np.float32)
    log(f"[setup] config {cfg} ({c['name']}): {p.shape[0]} particles + {g.shape[0]} ghosts "
        f"({time.time() - t0:.1f}s)")
    workers = ingest.default_workers()
    t1 = time.time()
    hb = ingest.make_block(allp, p.shape[0], (0, 0, 0), (1, 1, 1), workers=workers,
                           ghost_src=src, log=log, closed_first=p.shape[0])
    log(f"[setup] Delaunay: {hb.num_tets} tets with {workers} workers "
        f"({time.time() - t1:.1f}s)")
    return c, hb


def cpu_baseline(hb, grid, target_s):
    """The oracle (scalar C restatement of the reference path) on the first M
    cells of the same block, scaled to the whole block."""
    from oracle import oracle as O
    O.build()
    ob_full = O.Block.from_host(hb)
    n = hb.num_orig
    T = hb.num_tets

    def timed(m):
        sub = O.Block(hb.particles, hb.tets, hb.vert_to_tet, m, hb.bounds_min, hb.bounds_max)
        mt = max(1, int(T * m / n))
        t0 = time.perf_counter()
        cc = O.circumcenters(hb.tets[:mt], hb.particles)
        t1 = time.perf_counter()
        O.volumes(sub)
        O.dense_single(sub, grid)
        t2 = time.perf_counter()
        return (t1 - t0) * T / mt + (t2 - t1) * n / m, t2 - t0

    m = min(n, 20000)
    est, spent = timed(m)
    per_cell = spent / m
    m2 = int(min(n, max(m, target_s / max(per_cell, 1e-9))))
    if m2 > m:
        est, spent = timed(m2)
        m = m2
    del ob_full
    G = int(np.prod(grid))
    return {"value": G / est, "unit": "grid pts/s", "cores": 1, "kind": "port",
            "sample": f"first {m} of {n} cells ({spent:.1f}s of CPU work: circumcenters of "
                      f"the same fraction of tets, volume() and IterateCells over those "
                      f"cells), scaled to all cells; single thread",
            "est_full_s": est}


def exact_paths(run, grid, G, args):
    """The exact-Voronoi assignment in the mixed FP32/FP64 path and in the
    all-FP64 path on the same staged block: timing, FP64 fallbacks, and the
    agreement of the two (owner of every grid point, densities)."""
    from tess2_amd import host
    res = {}
    outs = {}
    for name, prec in (("mixed", host.MIXED), ("fp64", host.FP64)):
        for _ in range(args.warmup):
            run.run(grid, assignment=host.EXACT_VORONOI, precision=prec, want_volumes=False)
            run.sync()
        run.kernel_time()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            run.run(grid, assignment=host.EXACT_VORONOI, precision=prec, want_volumes=False)
            st = run.sync()
        ms = (time.perf_counter() - t0) * 1e3 / args.steps
        kms, _ = run.kernel_time()
        outs[name] = (run.owner(0, G), run.fetch(0, ngrid=G)["density"])
        res[name] = {"ms_per_step": ms, "grid_pts_per_s": G / (ms * 1e-3),
                     "owner_kernel_ms": kms, "fp64_fallbacks": int(st["fp64_fallbacks"]),
                     "cells_deposited": int(st["cells_scan"] + st["cells_cic"])}
        log(f"[exact {name}] {ms:.3f} ms/step, owner kernel {kms:.3f} ms, "
            f"fallbacks {st['fp64_fallbacks']}")
    om, dm = outs["mixed"]
    of, df = outs["fp64"]
    nz = df != 0
    rel = np.abs(dm[nz].astype(np.float64) - df[nz]) / np.abs(df[nz].astype(np.float64))
    res["speedup_mixed_vs_fp64"] = res["fp64"]["ms_per_step"] / res["mixed"]["ms_per_step"]
    res["owner_identical"] = bool(np.array_equal(om, of))
    res["max_rel_density_diff"] = float(rel.max()) if rel.size else 0.0
    res["rms_rel_density_diff"] = float(np.sqrt(np.mean(rel ** 2))) if rel.size else 0.0
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-exact", action="store_true",
                    help="skip the exact-Voronoi mixed / FP64 measurement")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--save-block", default=None,
                    help="tessellate, write the block to this .npz and exit")
    ap.add_argument("--load-block", default=None,
                    help="read the block written by --save-block instead of tessellating")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); this build has no "
              f"multi-GPU path", file=sys.stderr)
        sys.exit(2)

    from tess2_amd import host
    from tess2_amd.ingest import HostBlock  # noqa: F401

    c, hb = setup(args.config, args.load_block)
    if args.save_block:
        np.savez(args.save_block, particles=hb.particles, tets=hb.tets,
                 vert_to_tet=hb.vert_to_tet, num_orig=hb.num_orig,
                 bounds_min=hb.bounds_min, bounds_max=hb.bounds_max)
        log(f"[setup] wrote {args.save_block}")
        return
    gsz = c["grid"]
    grid = (gsz, gsz, gsz)
    G = gsz ** 3
    t0 = time.time()
    run = host.StagedRun([hb])
    log(f"[stage] H2D upload {time.time() - t0:.2f}s")
    for _ in range(args.warmup):
        run.run(grid)
        run.sync()
    run.kernel_time()               # reset the per-kernel event accumulator
    log(f"[run] warmup {args.warmup} done; timing {args.steps} steps")
    dev = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        run.run(grid)
        st = run.sync()
        dev.append(st["device_ms"])
    t1 = time.perf_counter()
    kern_ms, nk = run.kernel_time()
    ms = (t1 - t0) * 1e3 / args.steps
    log(f"[run] {ms:.3f} ms/step wall, {np.mean(dev):.3f} ms device, cell kernel "
        f"{kern_ms:.3f} ms avg over {nk}")
    N_all, N, T = hb.num_particles, hb.num_orig, hb.num_tets
    b_alg = 16 * N_all + 32 * T + 4 * N + 4 * G
    achieved = b_alg / (kern_ms * 1e-3) / 1e9
    traffic = None
    tf = os.path.join(ROOT, "profiles", f"traffic_{c['name']}.json")
    if os.path.exists(tf):
        traffic = json.load(open(tf)).get("hbm_bytes_per_launch")
    out = {
        "metric": METRIC, "value": G / (ms * 1e-3), "unit": "grid pts/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": c["name"], "particles": N, "ghosts": N_all - N, "tets": T,
                   "grid": list(grid), "assignment": "reference_scan",
                   "estimator": "DENSE_TESS", "device_ms_per_step": float(np.mean(dev)),
                   "cells_deposited": int(st["cells_scan"] + st["cells_cic"]),
                   "cells_cic": int(st["cells_cic"])},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "k_cells_ref", "kernel_ms": kern_ms, "alg_bytes": b_alg},
    }
    if not args.no_exact:
        out["exact_voronoi"] = exact_paths(run, grid, G, args)
    if not args.no_cpu:
        log("[cpu] timing the oracle on a sample")
        out["cpu_baseline"] = cpu_baseline(hb, grid, args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
