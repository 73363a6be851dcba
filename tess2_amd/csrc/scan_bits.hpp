// scan_bits.hpp -- one scan line of the reference's cell scan from the
// line's PtInCell verdicts held as bits.  Header-only and free of device
// intrinsics, so that a host build can check it against the point-by-point
// walk.
#pragma once

namespace t2 {

// The reference's two x walks along one scan line (src/dense.cpp:1546-1620)
// from the line's verdict bits (bit x = PtInCell at x, n0 <= 64 points): each
// walk moves one point at a time until the verdict changes, so it ends at the
// nearest bit of the other value, found here with one bit scan.
//   left walk from x0 = x_left: inside -> down through the run (x_left = the
//     first point outside, unchanged when the run reaches 0); outside -> up
//     to the first inside point (x_left = it)
//   right walk from x_right: x0 inside -> up through the run if x_right is
//     inside (x_right = the first point outside); x0 outside -> down to the
//     first inside point at or below x_right (x_right = it)
__host__ __device__ __forceinline__ void walk_row_bits(unsigned long long row, int n0, int &x_left,
                                              int &x_right, int &min_xi, int &max_xi)
{
    const unsigned long long all = n0 >= 64 ? ~0ull : (1ull << n0) - 1ull;
    const int x0 = x_left, xr = x_right;
    const bool x_in = (row >> x0) & 1ull;
    if (x_in) {
        const unsigned long long out = ~row & ((1ull << x0) - 1ull);   // outside, below x0
        int lo = 0;
        if (out) {
            lo = 64 - __builtin_clzll(out);   // one above the highest outside point
            x_left = lo - 1;
        }
        min_xi = lo;
        max_xi = x0;
        if ((row >> xr) & 1ull) {
            const unsigned long long outr = ~row & all & ~((2ull << xr) - 1ull);   // above xr
            int hi = n0 - 1;
            if (outr) {
                x_right = __builtin_ffsll(outr) - 1;
                hi = x_right - 1;
            }
            min_xi = min_xi < xr ? min_xi : xr;
            max_xi = max_xi > hi ? max_xi : hi;
        }
    } else {
        const unsigned long long inl = row & ~((2ull << x0) - 1ull);   // inside, above x0
        if (inl) {
            const int xa = __builtin_ffsll(inl) - 1;
            x_left = xa;
            min_xi = xa;
            max_xi = xa;
        }
        const unsigned long long inr = row & ((2ull << xr) - 1ull);    // inside, at or below xr
        if (inr) {
            const int xb = 63 - __builtin_clzll(inr);
            x_right = xb;
            min_xi = min_xi < xb ? min_xi : xb;
            max_xi = max_xi > xb ? max_xi : xb;
        }
    }
}

}  // namespace t2
