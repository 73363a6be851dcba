// scan_deposit.hpp -- the per-cell tail of the reference-exact density path,
// shared by the cell kernels: data-bounds filter and grid range of the cell
// box (src/dense.cpp:1381-1410), the scanline search (CellInteriorGridPts,
// src/dense.cpp:1475-1701) on PtInCell verdicts, and the deposit of mass/N
// per found point or, when the scan finds none, CIC of the site
// (src/dense.cpp:1437-1455,1692).  One 64-lane wavefront per cell; the
// cell's faces are held one per lane (face j*64 + lane in register j).
#pragma once

#include "deposit.hpp"
#include "geom.hpp"
#include "kernels.hpp"
#include "scan_bits.hpp"

namespace t2 {

enum { ST_SCAN = 0, ST_CIC = 1, ST_INCOMPLETE = 2, ST_OUTSIDE = 3, ST_DEGENERATE = 4 };

__device__ __forceinline__ bool wide_bit(unsigned long long w, int b) { return (w >> b) & 1ull; }

// PtInCell at (qx, qy, qz) against the F face planes staged in LDS: the
// largest and smallest significant distance decide it (a face with d >= 0
// exists iff hi >= 0, one with d < 0 iff lo < 0), the same float operations
// as a face-parallel probe's
__device__ __forceinline__ bool lds_verdict(float qx, float qy, float qz, int F,
                                            const float4 *s_fa, const float2 *s_fb, float eps)
{
    float hi = -INFINITY, lo = INFINITY;
#pragma unroll 4
    for (int k = 0; k < F; ++k) {
        const float4 a = s_fa[k];
        const float2 b = s_fb[k];
        const float d = a.x * (qx - a.w) + a.y * (qy - b.x) + a.z * (qz - b.y);
        const bool sig = fabsf(d) > eps;
        hi = fmaxf(hi, sig ? d : -INFINITY);
        lo = fminf(lo, sig ? d : INFINITY);
    }
    return !(hi >= 0.0f && lo < 0.0f);
}

// v: the cell; site: its particle; F faces with planes (fn*, ff*) per lane;
// mn*/mx*: the cell bbox, wave-uniform (the reference's fold, NaN included);
// s_fa / s_fb: LDS room for F face planes {n, f0.x} / {f0.y, f0.z}, already
// filled when faces_staged; runs: LDS room for RMAX scan runs.  WIDE adds
// the verdict masks of boxes up to 512 points, LINE the line windows of
// larger boxes (each costs registers: on only where such boxes are common).
template <int NF, int RMAX, bool WIDE, bool LINE>
__device__ __forceinline__ void scan_deposit(
    const int v, const float4 site, const int F, const float (&fnx)[NF], const float (&fny)[NF],
    const float (&fnz)[NF], const float (&ffx)[NF], const float (&ffy)[NF],
    const float (&ffz)[NF], float mnx, float mny, float mnz, float mxx, float mxy, float mxz,
    const GridP &gp, const CellOut &o, const Router *rt, const int self, float4 *s_fa,
    float2 *s_fb, const bool faces_staged, int4 *runs)
{
    const int lane = (int)(threadIdx.x & 63);
    const float ex = (gp.dmax[0] - gp.dmin[0]) * 2.0f * FLT_EPSILON;
    const float ey = (gp.dmax[1] - gp.dmin[1]) * 2.0f * FLT_EPSILON;
    const float ez = (gp.dmax[2] - gp.dmin[2]) * 2.0f * FLT_EPSILON;
    if (mnx < gp.dmin[0] - ex || mny < gp.dmin[1] - ey || mnz < gp.dmin[2] - ez ||
        mxx > gp.dmax[0] + ex || mxy > gp.dmax[1] + ey || mxz > gp.dmax[2] + ez) {
        if (lane == 0) {
            o.status[v] = ST_OUTSIDE;
            o.npts[v] = 0;
        }
        return;
    }
    if (isnan(mnx) || isnan(mny) || isnan(mnz) || isnan(mxx) || isnan(mxy) || isnan(mxz)) {
        if (lane == 0) {
            o.status[v] = ST_DEGENERATE;
            o.npts[v] = 0;
        }
        return;
    }
    const int cmi0 = phys2idx1(mnx, gp.step[0], gp.gmin[0]);
    const int cmi1 = phys2idx1(mny, gp.step[1], gp.gmin[1]);
    const int cmi2 = phys2idx1(mnz, gp.step[2], gp.gmin[2]);
    // every lane holds the same reduced bbox; the box dimensions are
    // made wave-uniform so that the scan's control runs on the scalar unit
    const int n0 = __builtin_amdgcn_readfirstlane(phys2idx1(mxx, gp.step[0], gp.gmin[0]) - cmi0 + 1);
    const int n1 = __builtin_amdgcn_readfirstlane(phys2idx1(mxy, gp.step[1], gp.gmin[1]) - cmi1 + 1);
    const int n2 = __builtin_amdgcn_readfirstlane(phys2idx1(mxz, gp.step[2], gp.gmin[2]) - cmi2 + 1);
    const float p0 = idx2phys1(cmi0, gp.step[0], gp.gmin[0]);
    const float p1 = idx2phys1(cmi1, gp.step[1], gp.gmin[1]);
    const float p2 = idx2phys1(cmi2, gp.step[2], gp.gmin[2]);
    const float eps = gp.eps;

    // PtInCell for the whole cell box at once when it has at most 64
    // points with each axis padded to a power of two (x to at most 32):
    // one lane per point, the faces broadcast from LDS, the verdicts
    // ballotted into one wave-uniform mask.  A verdict is the same function
    // of the same float operations as a face-parallel probe's; the scan
    // then walks each line with bit scans and reads single bits where it
    // probes.  Larger boxes: the wide and line-window masks below (WIDE /
    // LINE kernels), else face-parallel probes as the walk goes.
    const int lb0 = n0 > 1 ? 32 - __clz(n0 - 1) : 0;
    const int lb1 = n1 > 1 ? 32 - __clz(n1 - 1) : 0;
    const int lb2 = n2 > 1 ? 32 - __clz(n2 - 1) : 0;
    const int lb01 = lb0 + lb1;
    const bool use_mask = lb0 <= 5 && lb01 + lb2 <= 6;
    // Wider boxes (up to 512 points, lines of at most 32) in the layout
    // x + n0 (y + n1 z), 64 points per mask word (the batched kernel only:
    // its sparse blocks have such boxes, the dense single-block inputs
    // hardly any)
    const int nbox = n0 * n1 * n2;
    const bool use_wide = WIDE && !use_mask && n0 <= 32 && nbox <= 512;
    // Any larger box with lines of at most 64 points (LINE kernels): the
    // verdicts of a window of 64 >> lb0 lines of one z plane at once (lane
    // x + (line << lb0)) in one mask word, the next window when the walk
    // leaves it.  An empty line then costs one window share instead of the
    // ~2 n0 face-parallel probes of the two x walks.
    const bool use_line = LINE && !use_mask && !use_wide && n0 <= 64;
    unsigned long long mk = 0ull;
    // the wide masks' words: lane c holds word c (low / high half), read at
    // a wave-uniform lane index (no select chain over eight scalar words)
    unsigned wlo = 0u, whi = 0u;
    if ((use_mask || use_wide || use_line) && !faces_staged) {
        __syncthreads();
#pragma unroll
        for (int j = 0; j < NF; ++j) {
            const int f = j * 64 + lane;
            if (f < F) {
                s_fa[f] = make_float4(fnx[j], fny[j], fnz[j], ffx[j]);
                s_fb[f] = make_float2(ffy[j], ffz[j]);
            }
        }
        __syncthreads();
    }
    if (use_wide) {
        // point p = x + n0 (y + n1 z): quotients by the float reciprocal
        // of the divisor, exact for p < 512 (the error stays far below
        // the 0.5 / n margin of p + 0.5)
        const int n01 = n0 * n1;
        const float r01 = 1.0f / (float)n01, r0 = 1.0f / (float)n0;
        const int nw = (nbox + 63) >> 6;
        for (int c = 0; c < nw; ++c) {
            const int p = c * 64 + lane;
            const int z = (int)(((float)p + 0.5f) * r01);
            const int rem = p - z * n01;
            const int y = (int)(((float)rem + 0.5f) * r0);
            const int x = rem - y * n0;
            const float qx = p0 + (float)x * gp.step[0];
            const float qy = p1 + (float)y * gp.step[1];
            const float qz = p2 + (float)z * gp.step[2];
            const unsigned long long m =
                __ballot(p < nbox && lds_verdict(qx, qy, qz, F, s_fa, s_fb, eps));
            if (lane == c) {
                wlo = (unsigned)m;
                whi = (unsigned)(m >> 32);
            }
        }
    }
    if (use_mask) {
        const int x = lane & ((1 << lb0) - 1), y = (lane >> lb0) & ((1 << lb1) - 1);
        const int z = lane >> lb01;
        const float qx = p0 + (float)x * gp.step[0];
        const float qy = p1 + (float)y * gp.step[1];
        const float qz = p2 + (float)z * gp.step[2];
        const bool in_cell = lds_verdict(qx, qy, qz, F, s_fa, s_fb, eps);
        mk = __ballot(x < n0 && y < n1 && z < n2 && in_cell);
    }

    // PtInCell over all faces at once
    auto inside = [&](float qx, float qy, float qz) -> bool {
        bool pos = false, neg = false;
#pragma unroll
        for (int j = 0; j < NF; ++j) {
            if (j * 64 + lane < F) {
                const float d = fnx[j] * (qx - ffx[j]) + fny[j] * (qy - ffy[j]) +
                                fnz[j] * (qz - ffz[j]);
                if (fabsf(d) > eps) {
                    if (d >= 0.0f) pos = true;
                    else neg = true;
                }
            }
        }
        return !(__ballot(pos) != 0ull && __ballot(neg) != 0ull);
    };

#define WORD(k)                                                                               \
(((unsigned long long)(unsigned)__builtin_amdgcn_readlane((int)whi, (k)) << 32) |             \
 (unsigned)__builtin_amdgcn_readlane((int)wlo, (k)))
#define PROBE(xi, yi, zi)                                                                     \
(use_mask ? (bool)((mk >> ((xi) | ((yi) << lb0) | ((zi) << lb01))) & 1ull)                 \
 : use_wide ? wide_bit(WORD(((xi) + n0 * ((yi) + n1 * (zi))) >> 6),                        \
                       ((xi) + n0 * ((yi) + n1 * (zi))) & 63)                               \
          : inside(p0 + (float)(xi) * gp.step[0], p1 + (float)(yi) * gp.step[1],             \
                   p2 + (float)(zi) * gp.step[2]))

    // the points of runs[0, nb) (btot in all, run k's first at runs[k].w),
    // 64 at a time: one lane per point, which finds its run in LDS
    float mpp = 0.0f;
    auto deposit_runs = [&](const int nb, const int btot) {
        for (int base = 0; base < btot; base += 64) {
            const int p = base + lane;
            int r = 0;
            for (int k = 1; k < nb; ++k)
                if (p >= runs[k].w) r = k;
            if (p < btot) {
                const int4 rr = runs[r];
                deposit_point(cmi0 + rr.z + (p - rr.w), cmi1 + rr.y, cmi2 + rr.x, mpp,
                              cell_key(v, 0), gp, o, rt, self);
            }
        }
    };

    // ---- 5. scanline search (src/dense.cpp:1486-1676) --------------------
    // Pass 0 counts the points and keeps the runs (one per z-y line) in
    // LDS.  A cell with more than RMAX runs scans again in pass 1 -- the
    // state machine takes the same path -- now that mass/N is known, and
    // deposits its runs as it finds them: RMAX at a time as the buffer fills
    // in the LINE kernels, whose sparse inputs have many such cells, one at
    // a time in the others (a cell adds at most one contribution to a grid
    // point, so the order of its deposits does not matter).
    int nruns = 0, tot = 0;
    int nb = 0, btot = 0;                 // pass 1: runs buffered, their points
    int lz = -1, ly0 = 0;                 // the line window: plane, first line
    unsigned long long lw = 0ull;         // and its verdicts
    const int lwin = 64 >> lb0;           // lines per window (lb0 <= 6 there)
    for (int pass = 0; pass < 2; ++pass) {
        if (pass == 1) {
            if (nruns <= RMAX || tot == 0) break;
            mpp = gp.mass / (float)tot;   // src/dense.cpp:1692
        }
        int x_left = n0 / 2, x_right = n0 / 2, y_start = 0, first_x = 0;
        for (int zi = 0; zi < n2; zi++) {
            bool border_found = false, z_step_done = false;
            for (int yi = y_start; yi < n1; yi++) {
                int min_xi = n0 - 1, max_xi = 0;
                if (use_mask) {
                    const int base = (yi << lb0) | (zi << lb01);
                    const unsigned long long row = (mk >> base) & ((1ull << n0) - 1ull);
                    walk_row_bits(row, n0, x_left, x_right, min_xi, max_xi);
                } else if (use_wide) {
                    // the line's n0 <= 32 bits, from one word or two
                    const int base = n0 * (yi + n1 * zi), k = base >> 6, off = base & 63;
                    unsigned long long row = WORD(k) >> off;
                    if (off + n0 > 64) row |= WORD(k + 1) << (64 - off);
                    row &= (1ull << n0) - 1ull;
                    walk_row_bits(row, n0, x_left, x_right, min_xi, max_xi);
                } else if (use_line) {
                    if (zi != lz || yi < ly0 || yi >= ly0 + lwin) {
                        lz = zi;
                        ly0 = yi;
                        const int x = lane & ((1 << lb0) - 1), y = yi + (lane >> lb0);
                        const bool in_cell =
                            lds_verdict(p0 + (float)x * gp.step[0], p1 + (float)y * gp.step[1],
                                        p2 + (float)zi * gp.step[2], F, s_fa, s_fb, eps);
                        lw = __ballot(x < n0 && y < n1 && in_cell);
                    }
                    unsigned long long row = lw >> ((yi - ly0) << lb0);
                    if (n0 < 64) row &= (1ull << n0) - 1ull;
                    walk_row_bits(row, n0, x_left, x_right, min_xi, max_xi);
                } else {
                // the left walk starts where x_in was taken: its first
                // verdict is x_in (PtInCell is pure), each later one is
                // taken after the step
                const int x0 = x_left;
                const bool x_in = PROBE(x0, yi, zi);
                {
                    int xi = x0;
                    bool in_xi = x_in;
                    while (xi >= 0 && xi < n0) {
                        if (in_xi) {
                            if (xi < min_xi) min_xi = xi;
                            if (xi > max_xi) max_xi = xi;
                            if (x_in) xi--;
                            else {
                                x_left = xi;
                                break;
                            }
                        } else {
                            if (!x_in) xi++;
                            else {
                                x_left = xi;
                                break;
                            }
                        }
                        if (xi < 0 || xi >= n0) break;
                        in_xi = PROBE(xi, yi, zi);
                    }
                }
                for (int xi = x_right; xi >= 0 && xi < n0;) {
                    if (PROBE(xi, yi, zi)) {
                        if (xi < min_xi) min_xi = xi;
                        if (xi > max_xi) max_xi = xi;
                        if (x_in) xi++;
                        else {
                            x_right = xi;
                            break;
                        }
                    } else {
                        if (!x_in) xi--;
                        else {
                            x_right = xi;
                            break;
                        }
                    }
                }
                }
                int first_y = y_start;
                if (min_xi <= max_xi) {
                    if (pass == 0) {
                        if (nruns < RMAX && lane == 0)
                            runs[nruns] = make_int4(zi, yi, min_xi, tot);
                        nruns++;
                        tot += (max_xi - min_xi + 1);
                    } else if (LINE) {
                        if (nb == RMAX) {
                            __syncthreads();
                            deposit_runs(nb, btot);
                            __syncthreads();
                            nb = 0;
                            btot = 0;
                        }
                        if (lane == 0) runs[nb] = make_int4(zi, yi, min_xi, btot);
                        nb++;
                        btot += max_xi - min_xi + 1;
                    } else {
                        // one run at a time: the buffered deposit inside the
                        // walk costs these kernels registers, and their inputs
                        // rarely have cells of more than RMAX runs
                        const int len = max_xi - min_xi + 1;
                        for (int b0 = 0; b0 < len; b0 += 64)
                            if (b0 + lane < len)
                                deposit_point(cmi0 + min_xi + b0 + lane, cmi1 + yi, cmi2 + zi,
                                              mpp, cell_key(v, 0), gp, o, rt, self);
                    }
                    if (yi == y_start) first_x = (min_xi + max_xi) / 2;
                    if (!border_found) {
                        first_y = yi;
                        border_found = true;
                    }
                } else if (border_found) {
                    z_step_done = true;
                }
                if ((yi == n1 - 1 || z_step_done) && zi + 1 < n2) {
                    int yj;
                    for (yj = first_y; yj > 0; yj--) {
                        if (!PROBE(first_x, yj, zi + 1)) break;
                    }
                    y_start = yj;
                }
                if (z_step_done) break;
            }
        }
    }
#undef PROBE
#undef WORD
    __syncthreads();

    // ---- 6. deposit -------------------------------------------------------
    if (tot > 0 && nruns > RMAX) {
        if (LINE) deposit_runs(nb, btot);   // pass 1's last runs
        if (lane == 0) {
            o.status[v] = ST_SCAN;
            o.npts[v] = tot;
        }
    } else if (tot > 0) {
        mpp = gp.mass / (float)tot;   // src/dense.cpp:1692
        deposit_runs(nruns, tot);
        if (lane == 0) {
            o.status[v] = ST_SCAN;
            o.npts[v] = tot;
        }
    } else {
        int mn[3];
        float m8[8];
        cic_weights(site.x, site.y, site.z, gp.mass, gp, mn, m8);
        if (lane < 8) {
            float m = m8[0];
#pragma unroll
            for (int k = 1; k < 8; ++k)
                if (lane == k) m = m8[k];
            deposit_point(mn[0] + (lane & 1), mn[1] + ((lane >> 1) & 1), mn[2] + (lane >> 2),
                          m, cell_key(v, gp.project ? (lane >> 2) : 0), gp, o, rt, self);
        }
        if (lane == 0) {
            o.status[v] = ST_CIC;
            o.npts[v] = 8;
        }
    }
}

}  // namespace t2
