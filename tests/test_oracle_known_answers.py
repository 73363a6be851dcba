"""Hand-worked known answers for the parts of the oracle that no closed form
fixes: the CellInteriorGridPts scan state machine (src/dense.cpp:1475-1701)
and the first-seen (u, ut) order of neighbor_edges with the edge-ring start
of fill_edge_link (src/tet.cpp:164-270,389-409).

Every expected value below was traced by hand through the reference source,
step by step; the traces are written next to each case.  The scan cases run
the oracle's state machine on given PtInCell verdict masks (every probe of the
reference's scan is at a grid point cmin + i*step of the cell's box, so the
scan is a function of the box's verdicts by index; the oracle's geometric scan
calls the same walk with PtInCell as the verdict).

The mutation tests plant plausible mistakes in a copy of the oracle source
(compiled into a temporary directory, on the CPU) and check that these known
answers reject each one: x_in_right taken at x_right, a line's run starting
empty (min > max) instead of at (nx - 1, 0), x_left moved to 0 when
the left walk runs off the box, first_x taken from the z step's first found
line instead of the y_start line, no early z-step exit, the y_start probe at
x_left instead of first_x, the BFS visiting a tet's slots in reverse, an edge
ring circulating the other way, and an edge ring starting at another tet.
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

ORACLE_SRC = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                          "oracle")


def _mask(nx, ny, nz, inside):
    """inside: {(z, y): iterable of x} -> uint8 mask [z][y][x]."""
    m = np.zeros((nz, ny, nx), np.uint8)
    for (z, y), xs in inside.items():
        for x in xs:
            m[z, y, x] = 1
    return m


# Each case: (name, mask, expected total, expected runs [z][y] = (min_xi, max_xi)).
# Empty lines the scan visited keep the initial (nx-1, 0); lines it skipped
# (before y_start, after an early z-step exit) get (1, 0).
SCAN_CASES = [
    # One line, nx=8, inside x = 5..7.  x_left = x_right = 4, outside, so
    # both walks start with x_in = false.  Left walk: 4 out -> step right to
    # 5, in: record 5, x_left = 5, stop.  Right walk: 4 out -> step left
    # 3, 2, 1, 0 out, -1 ends the loop.  Only 5 is found, 6 and 7 are lost.
    ("run_right_of_x_left", _mask(8, 1, 1, {(0, 0): [5, 6, 7]}), 1, [[(5, 5)]]),

    # One line, nx=8, inside x = 1,2 and 5,6,7 (not an interval).  Left
    # walk: 4 out -> 5 in: record 5, x_left = 5.  Right walk: 4 out -> 3
    # out -> 2 in: record 2, x_right = 2.  The run is [min, max] = [2, 5]:
    # four points, two of them (3, 4) outside the verdicts.
    ("two_runs_joined", _mask(8, 1, 1, {(0, 0): [1, 2, 5, 6, 7]}), 4, [[(2, 5)]]),

    # nx=6, two lines.  y0: inside 0..3; x_left = x_right = 3, in.  Left:
    # 3, 2, 1, 0 in, -1 ends the loop without a break, so x_left stays 3.
    # Right: 3 in, 4 out -> x_right = 4.  Run (0, 3); first_x = 1.
    # y1: inside 0, 1; the probe at x_left = 3 is out.  Left: 3, 4, 5 out,
    # 6 ends.  Right from 4: 4, 3, 2 out, 1 in -> record 1, x_right = 1.
    # Run (1, 1): point 0 is lost.
    ("left_walk_reaches_0", _mask(6, 2, 1, {(0, 0): [0, 1, 2, 3], (0, 1): [0, 1]}),
     5, [[(0, 3), (1, 1)]]),

    # nx=4, ny=4.  y0 empty: both walks run off, (3, 0).  y1: inside 1, 2;
    # probe at x_left = 2 in.  Left: 2, 1 in, 0 out -> x_left = 0.  Right:
    # 2 in, 3 out -> x_right = 3.  Run (1, 2), border_found; y1 is not
    # y_start (0), so first_x stays 0.  y2 empty: probe at 0 out, left 0..3
    # out, right 3..0 out: (3, 0) and, with border_found, z_step_done: y3 is
    # filled (1, 0) and never scanned, so its points 1, 2 are lost.
    ("z_step_done_exit", _mask(4, 4, 1, {(0, 1): [1, 2], (0, 3): [1, 2]}),
     2, [[(3, 0), (1, 2), (3, 0), (1, 0)]]),

    # nx=7, two lines.  y0: inside 1..4; probe at 3 in.  Left: 3, 2, 1 in,
    # 0 out -> x_left = 0.  Right: 3, 4 in, 5 out -> x_right = 5.  Run (1, 4).
    # y1: inside 0..3.  Both flags come from x_left = 0: in.  Left: 0 in, -1
    # ends.  Right from x_right = 5 with x_in_right = true: 5 is out, which
    # reads as an exterior crossing, so x_right = 5 and the walk stops.
    # Run (0, 0): points 1, 2, 3 are lost.  (With the flag taken at x_right
    # the walk would step left from 5 and find 3.)
    ("x_in_right_from_x_left", _mask(7, 2, 1, {(0, 0): [1, 2, 3, 4], (0, 1): [0, 1, 2, 3]}),
     5, [[(1, 4), (0, 0)]]),

    # nx=ny=nz=3.  z0: only y2 = {0,1,2}.  y0, y1 empty (2, 0).  y2: probe
    # at 1 in; left 1, 0 in, -1 ends (x_left stays 1); right 1, 2 in, 3 ends
    # (x_right stays 1).  Run (0, 2), first found line, but not y_start, so
    # first_x stays 0; first_y = 2.  Last line: loosen y_start by probing
    # (first_x=0, y, z1) down from y=2: (0,2,z1) in, (0,1,z1) out -> y_start = 1.
    # z1 = {y0: 0..2, y1: 1,2, y2: 0,1}.  y0 skipped (1, 0), its points lost.
    # y1: probe at x_left = 1 in; left 1 in, 0 out -> x_left = 0; right 1, 2
    # in, 3 ends.  Run (1, 2); y1 == y_start -> first_x = 1.  y2: probe at 0
    # in; left 0 in, -1 ends; right from 1: 1 in, 2 out -> x_right = 2.  Run
    # (0, 1).  first_y = y_start = 1; last line: probe (1, y, z2) down from
    # y=1: (1,1,z2) in, loop ends at y=0 -> y_start = 0.
    # z2 = {y0: 1, y1: 0..2}.  y0: probe at x_left = 0 out; left 0 out, 1 in
    # -> record 1, x_left = 1; right from 2: out, 1 in -> record 1.  Run
    # (1, 1), first_x = 1.  y1: probe at 1 in; left 1, 0 in, -1 ends; right
    # 1, 2 in, 3 ends.  Run (0, 2).  y2 empty: probe at 1 out, left 1, 2 out,
    # right 1, 0 out: (2, 0), z_step_done; no z slice left.
    ("y_start_loosening",
     _mask(3, 3, 3, {(0, 2): [0, 1, 2],
                     (1, 0): [0, 1, 2], (1, 1): [1, 2], (1, 2): [0, 1],
                     (2, 0): [1], (2, 1): [0, 1, 2]}),
     11, [[(2, 0), (2, 0), (0, 2)], [(1, 0), (1, 2), (0, 1)], [(1, 1), (0, 2), (2, 0)]]),

    # As above but (0, 2, z1) is out: the loosening probe fails at once and
    # y_start = first_y = 2.  z1 = {y0: 0..2, y1: 1,2, y2: 1}: y0, y1 skipped
    # (1, 0).  y2: probe at x_left = 1 in; left 1 in, 0 out -> x_left = 0;
    # right 1 in, 2 out -> x_right = 2.  Run (1, 1); y2 == y_start -> first_x
    # = 1, first_y = 2.  Loosen: (1, 2, z2) out -> y_start = 2.
    # z2 = {y0: 1, y1: 0..2}: y0, y1 skipped; y2 empty: probe at 0 out, left
    # 0, 1, 2 out, right 2, 1, 0 out: (2, 0).
    ("y_start_stays_last_line",
     _mask(3, 3, 3, {(0, 2): [0, 1, 2],
                     (1, 0): [0, 1, 2], (1, 1): [1, 2], (1, 2): [1],
                     (2, 0): [1], (2, 1): [0, 1, 2]}),
     4, [[(2, 0), (2, 0), (0, 2)], [(1, 0), (1, 0), (1, 1)], [(1, 0), (1, 0), (2, 0)]]),

    # Box one point wide, 1 x 2 x 2, every verdict out.  Each line starts
    # with min_xi = nx - 1 = 0 and max_xi = 0; the probe at x_left = 0 is
    # out, the left walk steps to 1 and ends, the right walk steps to -1 and
    # ends, and min <= max still holds: the line keeps point 0.  z0: y0
    # (0, 0), first_x = 0, border_found; y1 (0, 0); last line: loosen from
    # first_y = y_start = 0 -> y_start = 0.  z1 the same.  Four points.
    ("one_point_wide_box", _mask(1, 2, 2, {}), 4, [[(0, 0), (0, 0)], [(0, 0), (0, 0)]]),

    # Empty box 3 x 2 x 2: every line visited, both walks run off: (2, 0);
    # no point found, so the cell falls back to CIC of its site
    # (src/dense.cpp:1437-1455).
    ("empty_box", _mask(3, 2, 2, {}), 0, [[(2, 0), (2, 0)], [(2, 0), (2, 0)]]),
]


@pytest.mark.parametrize("name,mask,tot,runs", SCAN_CASES, ids=[c[0] for c in SCAN_CASES])
def test_scan_known_answers(oracle, name, mask, tot, runs):
    got_tot, got_runs = oracle.scan_mask(mask)
    assert got_runs == runs
    assert got_tot == tot
    # the total is the sum of the run lengths (the deposit loop, :1680-1696)
    assert tot == sum(b - a + 1 for z in runs for a, b in z if a <= b)


# ---- star traversal and edge rings ----------------------------------------
# The octahedron around vertex 0 at the origin: 1 = +x, 2 = -x, 3 = +y,
# 4 = -y, 5 = +z, 6 = -z; one tet (0, a, b, c) per octant, a in {1,2},
# b in {3,4}, c in {5,6}, with vertex slots deliberately shuffled.
# tets[i] lies opposite verts[i]; the outer faces (opposite 0) are hull (-1).
OCTA_TETS = np.array([
    [3, 0, 5, 1, 2, -1, 4, 1],     # t0 (+x, +y, +z)
    [5, 2, 3, 0, 5, 0, 3, -1],     # t1 (-x, +y, +z)
    [0, 4, 1, 5, -1, 0, 3, 6],     # t2 (+x, -y, +z)
    [4, 5, 0, 2, 1, 7, -1, 2],     # t3 (-x, -y, +z)
    [6, 1, 0, 3, 0, 5, -1, 6],     # t4 (+x, +y, -z)
    [0, 3, 6, 2, -1, 7, 1, 4],     # t5 (-x, +y, -z)
    [1, 6, 4, 0, 7, 2, 4, -1],     # t6 (+x, -y, -z)
    [2, 0, 4, 6, 6, -1, 5, 3],     # t7 (-x, -y, -z)
], np.int32)
OCTA_POINTS = np.float32([[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0],
                          [0, 0, 1], [0, 0, -1]])

# BFS of vertex 0 from t7 (the last tet holding 0, as fill_vert_to_tet gives):
# pop t7 [2,0,4,6]: u=2 new (2,7), queue t6; 0 skipped; u=4 new (4,7), queue
# t5; u=6 new (6,7), queue t3.  Pop t6 [1,6,4,0]: u=1 new (1,6), queue t7, t2,
# t4.  Pop t5 [0,3,6,2]: u=3 new (3,5), queue t7, t1, t4.  Pop t3 [4,5,0,2]:
# u=5 new (5,3).  All later pops find no new vertex; no hull link is taken
# (each -1 sits in the slot of vertex 0), so the cell is finite.
OCTA_V0_NBRS = [(2, 7), (4, 7), (6, 7), (1, 6), (3, 5), (5, 3)]

# Edge rings of vertex 0, each from its first-seen tet ut: circulate_start
# takes the first slot holding neither 0 nor u; circulate_next crosses the
# face opposite that slot and continues with the vertex that is neither 0, u
# nor the crossed one.  E.g. (0, 5) from t3 [4,5,0,2]: slot 0 (vertex 4);
# cross t3.tets[0] = t1 carrying vertex 2 (slot 1 of t1); cross t1.tets[1] =
# t0 carrying 3 (slot 0); cross t0.tets[0] = t2 carrying 1 (slot 2); cross
# t2.tets[2] = t3 = ut, stop.
OCTA_V0_RINGS = {2: [7, 5, 1, 3], 4: [7, 6, 2, 3], 6: [7, 6, 4, 5],
                 1: [6, 2, 0, 4], 3: [5, 1, 0, 4], 5: [3, 1, 0, 2]}

# Vertex 5 (+z) lies on the hull.  BFS from t3 [4,5,0,2]: u=4 new (4,3),
# queue t1; 5 skipped; u=0 new (0,3), its link is -1 -> not finite; u=2 new
# (2,3), queue t2.  Pop t1 [5,2,3,0]: u=3 new (3,1).  Pop t2 [0,4,1,5]: u=1
# new (1,2).  Pop t0: nothing new.
OCTA_V5_NBRS = [(4, 3), (0, 3), (2, 3), (3, 1), (1, 2)]


def test_octahedron_star_order(oracle):
    fin, nbrs = oracle.neighbor_edges(0, OCTA_TETS, 7)
    assert fin == 1 and nbrs == OCTA_V0_NBRS
    assert oracle.complete(0, OCTA_TETS, 7) == 1
    fin, nbrs = oracle.neighbor_edges(5, OCTA_TETS, 3)
    assert fin == 0 and nbrs == OCTA_V5_NBRS
    assert oracle.complete(5, OCTA_TETS, 3) == 0


def test_octahedron_edge_rings(oracle):
    for u, ut in OCTA_V0_NBRS:
        assert oracle.edge_link(0, u, ut, OCTA_TETS) == OCTA_V0_RINGS[u]


def test_octahedron_cell_volume(oracle):
    """The Voronoi cell of the centre is the cube [-1/2, 1/2]^3: each face a
    unit square (two fan triangles of area 1/2, exact), each neighbour at
    distance 1, so volume() sums six float32 terms 1*1/6."""
    from oracle.oracle import Block
    vtt = np.array([7, 6, 7, 5, 7, 3, 7], np.int32)
    b = Block(OCTA_POINTS, OCTA_TETS, vtt, 1, [-1, -1, -1], [1, 1, 1])
    cc = oracle.circumcenters(OCTA_TETS, OCTA_POINTS)
    assert np.array_equal(np.abs(cc), np.full((8, 3), 0.5, np.float32))
    sixth = np.float32(1) / np.float32(6)
    want = np.float32(0)
    for _ in range(6):
        want = np.float32(want + sixth)
    assert oracle.volumes(b, cc)[0] == want


# ---- mutation checks: the known answers reject plausible mistakes ----------

MUTANTS = {
    "x_in_right_at_x_right": [(
        "x_in_left = x_in_right = inside(ctx, x_left, yi, zi) ? 1 : 0;",
        "x_in_left = inside(ctx, x_left, yi, zi) ? 1 : 0;"
        " x_in_right = inside(ctx, x_right, yi, zi) ? 1 : 0;")],
    "x_left_reset_at_0": [(
        "            /* x step right (:1585-1620) */",
        "            if (xi < 0) x_left = 0;\n            /* x step right (:1585-1620) */")],
    "first_x_from_first_found_line": [(
        "                if (yi == y_start)\n                    first_x",
        "                if (!border_found)\n                    first_x")],
    "no_early_z_exit": [(
        "            if (min_xi > max_xi && border_found)\n                z_step_done = 1;",
        "            if (min_xi > max_xi && border_found)\n                z_step_done = 0;")],
    "run_starts_empty": [(
        "            min_xi = cgp[0] - 1;\n            max_xi = 0;",
        "            min_xi = cgp[0];\n            max_xi = -1;")],
    "loosen_probe_at_x_left": [(
        "if (!inside(ctx, first_x, yj, zi + 1))",
        "if (!inside(ctx, x_left, yj, zi + 1))")],
    "bfs_slots_reversed": [(
        "        ib_push(vis, cur);\n        for (int i = 0; i < 4; ++i) {",
        "        ib_push(vis, cur);\n        for (int i = 3; i >= 0; --i) {")],
    "ring_other_direction": [(
        "    for (int i = 0; i < 4; ++i) {\n        int w = tets[t].verts[i];\n"
        "        if (w != x && w != y)\n            return i;",
        "    for (int i = 3; i >= 0; --i) {\n        int w = tets[t].verts[i];\n"
        "        if (w != x && w != y)\n            return i;")],
    "ring_starts_at_next_tet": [(
        "        t = nt;\n        wi = nwi;\n    }\n}",
        "        t = nt;\n        wi = nwi;\n    }\n"
        "    if (link->n > 1) {\n        int f0 = link->p[0];\n"
        "        memmove(link->p, link->p + 1, sizeof(int) * (size_t)(link->n - 1));\n"
        "        link->p[link->n - 1] = f0;\n    }\n}")],
}


def _build_mutant(tmp_path, name):
    src = open(os.path.join(ORACLE_SRC, "tess_oracle.c")).read()
    for old, new in MUTANTS[name]:
        assert src.count(old) == 1, f"mutation site of {name} not unique"
        src = src.replace(old, new)
    c = tmp_path / f"{name}.c"
    c.write_text(src)
    so = tmp_path / f"{name}.so"
    subprocess.check_call(["gcc", "-O1", "-w", "-fPIC", "-std=c99", "-ffp-contract=off",
                           "-I", ORACLE_SRC, "-shared", "-o", str(so), str(c), "-lm"])
    L = C.CDLL(str(so))
    I32P = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
    U8P = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
    L.o_scan_mask.argtypes = [I32P, U8P, I32P]
    L.o_neighbor_edges.argtypes = [C.c_int, I32P, C.c_int, I32P, I32P, C.c_int,
                                   C.POINTER(C.c_int)]
    L.o_fill_edge_link.argtypes = [I32P, C.c_int, C.c_int, C.c_int, C.c_int, I32P]
    return L


def _known_answers_hold(L):
    for _, mask, tot, runs in SCAN_CASES:
        nz, ny, nx = mask.shape
        border = np.zeros(2 * ny * nz, np.int32)
        got = L.o_scan_mask(np.array([nx, ny, nz], np.int32), mask.reshape(-1), border)
        b = border.reshape(nz, ny, 2)
        if got != tot or [[tuple(b[z, y]) for y in range(ny)] for z in range(nz)] != runs:
            return False
    tets = OCTA_TETS.reshape(-1)
    u = np.zeros(64, np.int32)
    ut = np.zeros(64, np.int32)
    n = C.c_int()
    L.o_neighbor_edges(0, tets, 7, u, ut, 64, C.byref(n))
    if list(zip(u[:n.value].tolist(), ut[:n.value].tolist())) != OCTA_V0_NBRS:
        return False
    for uu, t in OCTA_V0_NBRS:
        link = np.zeros(64, np.int32)
        k = L.o_fill_edge_link(link, 64, 0, uu, t, tets)
        if link[:k].tolist() != OCTA_V0_RINGS[uu]:
            return False
    return True


def test_unmutated_copy_passes(tmp_path):
    src = open(os.path.join(ORACLE_SRC, "tess_oracle.c")).read()
    (tmp_path / "plain.c").write_text(src)
    MUTANTS["_plain"] = []
    try:
        assert _known_answers_hold(_build_mutant(tmp_path, "_plain"))
    finally:
        del MUTANTS["_plain"]


@pytest.mark.parametrize("name", sorted(MUTANTS))
def test_known_answers_reject_mutant(tmp_path, name):
    assert not _known_answers_hold(_build_mutant(tmp_path, name))
