"""The cell scan's bit-scan line walk (tess2_amd/csrc/scan_bits.hpp) against
the reference's point-by-point x walks (src/dense.cpp:1546-1620), on the
host: every verdict pattern of lines up to 12 points with every start pair,
and 2M random lines of 13-64 points (with all-inside, all-outside and
single-point lines of 64 at both ends).  The header is compiled with g++ with
the HIP qualifiers defined away; the walk below is the reference's loop
with PtInCell replaced by a bit lookup."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "tess2_amd", "csrc")

HARNESS = r"""
#define __host__
#define __device__
#define __forceinline__ inline
#include <cstdio>
#include "scan_bits.hpp"

// src/dense.cpp:1546-1620 with PtInCell(x) = bit x of row
static void walk_points(unsigned long long row, int n0, int &x_left, int &x_right,
                        int &min_xi, int &max_xi)
{
    auto in = [&](int x) { return (bool)((row >> x) & 1ull); };
    const bool x_in = in(x_left);
    min_xi = n0 - 1;
    max_xi = 0;
    for (int xi = x_left; xi >= 0 && xi < n0;) {
        if (in(xi)) {
            if (xi < min_xi) min_xi = xi;
            if (xi > max_xi) max_xi = xi;
            if (x_in) xi--;
            else { x_left = xi; break; }
        } else {
            if (!x_in) xi++;
            else { x_left = xi; break; }
        }
    }
    for (int xi = x_right; xi >= 0 && xi < n0;) {
        if (in(xi)) {
            if (xi < min_xi) min_xi = xi;
            if (xi > max_xi) max_xi = xi;
            if (x_in) xi++;
            else { x_right = xi; break; }
        } else {
            if (!x_in) xi--;
            else { x_right = xi; break; }
        }
    }
}

static long bad = 0, cases = 0;
static void check(unsigned long long r, int n0, int xl, int xr)
{
    int a1 = xl, b1 = xr, mn1, mx1;
    int a2 = xl, b2 = xr, mn2 = n0 - 1, mx2 = 0;
    walk_points(r, n0, a1, b1, mn1, mx1);
    t2::walk_row_bits(r, n0, a2, b2, mn2, mx2);
    ++cases;
    if (a1 != a2 || b1 != b2 || mn1 != mn2 || mx1 != mx2) {
        if (bad++ < 10)
            printf("n0 %d row %llx x_left %d x_right %d: points %d %d %d %d, bits %d %d %d %d\n",
                   n0, r, xl, xr, a1, b1, mn1, mx1, a2, b2, mn2, mx2);
    }
}

int main()
{
    for (int n0 = 1; n0 <= 12; ++n0)
        for (unsigned long long r = 0; r < (1ull << n0); ++r)
            for (int xl = 0; xl < n0; ++xl)
                for (int xr = 0; xr < n0; ++xr) check(r, n0, xl, xr);
    unsigned long long s = 88172645463325252ull;   // xorshift64
    for (int it = 0; it < 2000000; ++it) {
        s ^= s << 13; s ^= s >> 7; s ^= s << 17;
        const int n0 = 13 + (int)(s % 52);
        const unsigned long long all = n0 == 64 ? ~0ull : (1ull << n0) - 1ull;
        unsigned long long r = (s >> 11) & all;
        if (it & 1) r |= (s >> 3) & all;          // denser lines
        if (it % 3 == 0) r = all & ~(r & (r >> 1)); // long runs
        check(r, n0, (int)((s >> 40) % n0), (int)((s >> 50) % n0));
    }
    const unsigned long long edge[] = {0ull, ~0ull, 1ull, 1ull << 63, 3ull << 62, 3ull,
                                       ~1ull, ~(1ull << 63), 0x00ffff0000ffff00ull};
    for (unsigned long long r : edge)
        for (int xl = 0; xl < 64; ++xl)
            for (int xr = 0; xr < 64; ++xr) check(r, 64, xl, xr);
    printf("cases %ld mismatches %ld\n", cases, bad);
    return bad != 0;
}
"""


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_bit_walk_matches_point_walk(tmp_path):
    src = tmp_path / "walk.cpp"
    exe = tmp_path / "walk"
    src.write_text(HARNESS)
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-I", CSRC, str(src), "-o", str(exe)])
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "mismatches 0" in out.stdout
