// exact.hip -- the exact-Voronoi assignment: every grid point goes to its
// nearest site (ties to the lowest particle id), located through a uniform
// bucket hash of the particles built on the device for each run.
//
// Precision modes:
//   MIXED: squared distances in float32; the FP32 decision stands only when
//          the runner-up is outside a relative window of 32 ulp around the
//          winner (each FP32 squared distance is within ~5 ulp of exact), so
//          the FP32 choice equals the exact and the FP64 choice; otherwise
//          the point is searched again in double (counted as a fallback).
//          Deposit decisions of cells near the data-bounds threshold, or with
//          ill-conditioned circumcenters, are re-made in double the same way.
//   FP64:  every distance, circumcenter and bound test in double.
#include "kernels.hpp"
#include "geom.hpp"
#include "deposit.hpp"

#include <algorithm>

namespace t2 {

enum { EX_DEPOSIT = 0, EX_INCOMPLETE = 2, EX_OUTSIDE = 3, EX_DEGENERATE = 4 };

// ---------------------------------------------------------------------------
// bucket hash: count, exclusive scan, scatter (particles sorted by bucket,
// stored as float4 {x, y, z, id bits})
// ---------------------------------------------------------------------------
__device__ __forceinline__ int bucket_of(float4 p, const HashP &h)
{
    int c[3];
    const float q[3] = {p.x, p.y, p.z};
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        int ci = (int)floorf((q[i] - h.lo[i]) * h.inv);
        ci = ci < 0 ? 0 : (ci >= h.dims[i] ? h.dims[i] - 1 : ci);
        c[i] = ci;
    }
    return (c[2] * h.dims[1] + c[1]) * h.dims[0] + c[0];
}

__global__ void k_hash_count(const float4 *__restrict__ pts, int np, HashP h,
                             unsigned *__restrict__ cnt, int *__restrict__ bk)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    const int b = bucket_of(pts[p], h);
    bk[p] = b;
    atomicAdd(&cnt[b], 1u);
}

// exclusive scan in three passes (1024 elements per block)
__global__ void k_scan_local(const unsigned *__restrict__ in, unsigned *__restrict__ out,
                             unsigned *__restrict__ sums, int n)
{
    __shared__ unsigned s[1024];
    const int i = blockIdx.x * 1024 + threadIdx.x;
    const unsigned x = i < n ? in[i] : 0u;
    s[threadIdx.x] = x;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const unsigned y = threadIdx.x >= (unsigned)off ? s[threadIdx.x - off] : 0u;
        __syncthreads();
        s[threadIdx.x] += y;
        __syncthreads();
    }
    if (i < n) out[i] = s[threadIdx.x] - x;
    if (threadIdx.x == 1023) sums[blockIdx.x] = s[1023];
}

__global__ void k_scan_sums(unsigned *__restrict__ sums, int nb)
{
    // one block of 1024 threads; nb <= 1024 * 16
    __shared__ unsigned s[1024];
    unsigned loc[16];
    unsigned run = 0;
    const int per = (nb + 1023) / 1024;
    for (int k = 0; k < per; ++k) {
        const int i = threadIdx.x * per + k;
        const unsigned x = i < nb ? sums[i] : 0u;
        loc[k] = run;
        run += x;
    }
    s[threadIdx.x] = run;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const unsigned y = threadIdx.x >= (unsigned)off ? s[threadIdx.x - off] : 0u;
        __syncthreads();
        s[threadIdx.x] += y;
        __syncthreads();
    }
    const unsigned base = s[threadIdx.x] - run;
    for (int k = 0; k < per; ++k) {
        const int i = threadIdx.x * per + k;
        if (i < nb) sums[i] = base + loc[k];
    }
}

__global__ void k_scan_add(unsigned *__restrict__ out, const unsigned *__restrict__ sums, int n,
                           unsigned *__restrict__ fill)
{
    const int i = blockIdx.x * 1024 + threadIdx.x;
    if (i < n) {
        const unsigned v = out[i] + sums[blockIdx.x];
        out[i] = v;
        fill[i] = v;
    }
}

// the fourth word is the particle's id: its index, or with several blocks
// its home key (rank << 23 | cell), which also breaks distance ties
__global__ void k_hash_scatter(const float4 *__restrict__ pts, int np, const int *__restrict__ bk,
                               unsigned *__restrict__ fill, float4 *__restrict__ sorted,
                               const int *__restrict__ key)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    const unsigned slot = atomicAdd(&fill[bk[p]], 1u);
    const float4 q = pts[p];
    sorted[slot] = make_float4(q.x, q.y, q.z, __int_as_float(key ? key[p] : p));
}

void launch_hash_build(const float4 *pts, int np, const HashP &h, unsigned *start,
                       unsigned *fill, unsigned *sums, int *bk, float4 *sorted, const int *key,
                       hipStream_t s)
{
    const int nbk = h.dims[0] * h.dims[1] * h.dims[2];
    launch_fill_u32(start, (size_t)nbk + 1, 0u, s);
    if (np > 0)
        hipLaunchKernelGGL(k_hash_count, dim3((np + 255) / 256), dim3(256), 0, s, pts, np, h,
                           start, bk);
    const int n = nbk + 1;   // the extra slot turns into the total
    const int nblk = (n + 1023) / 1024;
    hipLaunchKernelGGL(k_scan_local, dim3(nblk), dim3(1024), 0, s, start, start, sums, n);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(1024), 0, s, sums, nblk);
    hipLaunchKernelGGL(k_scan_add, dim3(nblk), dim3(1024), 0, s, start, sums, n, fill);
    if (np > 0)
        hipLaunchKernelGGL(k_hash_scatter, dim3((np + 255) / 256), dim3(256), 0, s, pts, np, bk,
                           fill, sorted, key);
}

// ---------------------------------------------------------------------------
// nearest site of a grid point
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool better64(double d, int id, double bd, int bid)
{
    return d < bd || (d == bd && id < bid);
}

// The buckets of shell r around bucket c (Chebyshev distance r), as runs of
// the bucket-sorted particle array: buckets are numbered x fastest, so the
// buckets of a row that lies on the shell's z or y faces form one run
// [start[b0], start[b1 + 1]); an inner row contributes its two end buckets.
// Each (dz, dy) row gives at most two runs, visited in order by the caller.
struct ShellRow {
    unsigned a0, a1, b0, b1;   // runs [a0, a1) and [b0, b1)
};

__device__ __forceinline__ ShellRow shell_row(int r, int dz, int dy, int y, int z, int cx,
                                              const HashP &h, const unsigned *__restrict__ start)
{
    const int row = (z * h.dims[1] + y) * h.dims[0];
    const bool face = (dz == -r || dz == r || dy == -r || dy == r);
    ShellRow s{0u, 0u, 0u, 0u};
    if (face) {
        const int xa = max(cx - r, 0), xb = min(cx + r, h.dims[0] - 1);
        s.a0 = start[row + xa];
        s.a1 = start[row + xb + 1];
    } else {
        if (cx - r >= 0) {
            s.a0 = start[row + cx - r];
            s.a1 = start[row + cx - r + 1];
        }
        if (cx + r < h.dims[0]) {
            s.b0 = start[row + cx + r];
            s.b1 = start[row + cx + r + 1];
        }
    }
    return s;
}

__device__ __forceinline__ void home_bucket(float gx, float gy, float gz, const HashP &h, int c[3])
{
    const float q[3] = {gx, gy, gz};
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        int ci = (int)floorf((q[i] - h.lo[i]) * h.inv);
        c[i] = ci < 0 ? 0 : (ci >= h.dims[i] ? h.dims[i] - 1 : ci);
    }
}

// search in double; returns the owner id
__device__ int nearest_f64(float gx, float gy, float gz, const HashP &h,
                           const unsigned *__restrict__ start, const float4 *__restrict__ hp)
{
    int c[3];
    home_bucket(gx, gy, gz, h, c);
    double best = 1e300;
    int bid = 0x7fffffff;
    const int rmax = h.dims[0] + h.dims[1] + h.dims[2];
    for (int r = 0; r <= rmax; ++r) {
        for (int dz = -r; dz <= r; ++dz) {
            const int z = c[2] + dz;
            if (z < 0 || z >= h.dims[2]) continue;
            for (int dy = -r; dy <= r; ++dy) {
                const int y = c[1] + dy;
                if (y < 0 || y >= h.dims[1]) continue;
                const ShellRow sr = shell_row(r, dz, dy, y, z, c[0], h, start);
                for (int k = 0; k < 2; ++k) {
                    const unsigned s1 = k ? sr.b1 : sr.a1;
                    for (unsigned s = k ? sr.b0 : sr.a0; s < s1; ++s) {
                        const float4 p = hp[s];
                        const double ex = (double)gx - (double)p.x;
                        const double ey = (double)gy - (double)p.y;
                        const double ez = (double)gz - (double)p.z;
                        const double d = ex * ex + ey * ey + ez * ez;
                        const int id = __float_as_int(p.w);
                        if (better64(d, id, best, bid)) {
                            best = d;
                            bid = id;
                        }
                    }
                }
            }
        }
        // every bucket beyond shell r is at least r*h away
        // (r - 0.001) cells: the bucket of a point near a bucket face can be
        // off by ~1e-4 of a cell through the float bucket arithmetic
        const double lb = ((double)r - 0.001) * (double)h.cell;
        if (bid != 0x7fffffff && r > 0 && lb * lb > best * (1.0 + 1e-9) + 1e-300) break;
    }
    return bid;
}

// search in float with the tie window; returns owner or -1 when undecided
__device__ int nearest_f32(float gx, float gy, float gz, const HashP &h,
                           const unsigned *__restrict__ start, const float4 *__restrict__ hp)
{
    int c[3];
    home_bucket(gx, gy, gz, h, c);
    float best = INFINITY, second = INFINITY;
    int bid = -1;
    const int rmax = h.dims[0] + h.dims[1] + h.dims[2];
    for (int r = 0; r <= rmax; ++r) {
        for (int dz = -r; dz <= r; ++dz) {
            const int z = c[2] + dz;
            if (z < 0 || z >= h.dims[2]) continue;
            for (int dy = -r; dy <= r; ++dy) {
                const int y = c[1] + dy;
                if (y < 0 || y >= h.dims[1]) continue;
                const ShellRow sr = shell_row(r, dz, dy, y, z, c[0], h, start);
                for (int k = 0; k < 2; ++k) {
                    const unsigned s1 = k ? sr.b1 : sr.a1;
                    for (unsigned s = k ? sr.b0 : sr.a0; s < s1; ++s) {
                        const float4 p = hp[s];
                        const float ex = gx - p.x, ey = gy - p.y, ez = gz - p.z;
                        const float d = ex * ex + ey * ey + ez * ez;
                        if (d < best) {
                            second = best;
                            best = d;
                            bid = __float_as_int(p.w);
                        } else if (d < second) {
                            second = d;
                        }
                    }
                }
            }
        }
        const float lb = ((float)r - 0.001f) * h.cell;
        if (bid >= 0 && r > 0 && lb * lb > best * 1.0001f + 1e-30f) break;
    }
    // decided only if the runner-up is clearly farther (32 ulp window) and
    // nothing underflowed
    const float tau = 32.0f * 5.9604645e-8f;
    if (bid >= 0 && best > 1e-30f && second > best * (1.0f + tau)) return bid;
    if (bid >= 0 && best == 0.0f && second > 1e-30f) {
        // the grid point coincides with a site: exact in float
        return bid;
    }
    return -1;
}

// one thread per block grid point; a 64-lane wave covers a 4x4x4 tile
// With several blocks the owner is a home key (rank << 23 | cell) and the
// counts go to the home block's cell (npts_tab / stex_tab by rank).
__global__ __launch_bounds__(256) void k_owner(GridP gp, HashP h,
                                               const unsigned *__restrict__ start,
                                               const float4 *__restrict__ hp, int *__restrict__ owner,
                                               int prec, unsigned *ctrs, int tiles_x, int tiles_y,
                                               int ntiles, const int *__restrict__ status,
                                               int *__restrict__ npts, int ncells,
                                               int *const *__restrict__ npts_tab,
                                               const int *const *__restrict__ stex_tab,
                                               unsigned *__restrict__ defq, unsigned defq_cap)
{
    const int wave = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (wave >= ntiles) return;
    const int lane = threadIdx.x & 63;
    const int tx = wave % tiles_x, ty = (wave / tiles_x) % tiles_y, tz = wave / (tiles_x * tiles_y);
    const int lx0 = tx * 4 + (lane & 3), ly0 = ty * 4 + ((lane >> 2) & 3), lz0 = tz * 4 + (lane >> 4);
    const bool on_grid = lx0 < gp.bnum_idx[0] && ly0 < gp.bnum_idx[1] && lz0 < gp.bnum_idx[2];
    // off-grid lanes of an edge tile stand in for an on-grid point of the
    // tile but write nothing and count nothing
    const int lx = on_grid ? lx0 : tx * 4, ly = on_grid ? ly0 : ty * 4, lz = on_grid ? lz0 : tz * 4;
    const long long lidx = ((long long)lz * gp.bnum_idx[1] + ly) * gp.bnum_idx[0] + lx;
    const int gx = gp.bmin_idx[0] + lx, gy = gp.bmin_idx[1] + ly, gz = gp.bmin_idx[2] + lz;
    const float px = idx2phys1(gx, gp.step[0], gp.gmin[0]);
    const float py = idx2phys1(gy, gp.step[1], gp.gmin[1]);
    const float pz = idx2phys1(gz, gp.step[2], gp.gmin[2]);
    const bool in_block = px >= gp.bmin[0] && px <= gp.bmax[0] && py >= gp.bmin[1] &&
                          py <= gp.bmax[1] && pz >= gp.bmin[2] && pz <= gp.bmax[2];
    if (!in_block && on_grid) owner[lidx] = -1;
    int o = -1;
    bool failed = false;
    if (in_block && on_grid) {
        if (prec == 0) o = nearest_f32(px, py, pz, h, start, hp);
        failed = o < 0 && prec == 0;
    }
    // FP32 filter failures go to a queue that k_owner_deferred searches in
    // FP64 with full waves: searched here, one failing lane would hold its
    // 63 neighbours for a whole FP64 search.  C_DEFER counts every failure
    // (one atomic per wave); beyond the queue's end the search runs inline.
    bool defer = false;
    const unsigned long long dm = __ballot(failed);
    if (dm) {
        const int leader = __ffsll(dm) - 1;
        unsigned base = 0;
        if (lane == leader) base = atomicAdd(&ctrs[C_DEFER], (unsigned)__popcll(dm));
        base = __shfl(base, leader);
        const unsigned slot = base + (unsigned)__popcll(dm & ((1ull << lane) - 1ull));
        if (failed && slot < defq_cap) {
            defq[slot] = (unsigned)lidx;
            defer = true;
        }
    }
    if (in_block && on_grid) {
        if (o < 0 && !defer) o = nearest_f64(px, py, pz, h, start, hp);
        owner[lidx] = o;           // -1 for a deferred point until its FP64 search
    }
    // points per deposited cell; the lanes of a tile mostly share owners, so
    // one atomic per distinct owner in the wave
    const int hc = npts_tab ? (o & 0x7fffff) : o;
    const int *hst = npts_tab && o >= 0 ? stex_tab[o >> 23] : status;
    int *hnp = npts_tab && o >= 0 ? npts_tab[o >> 23] : npts;
    const bool cnt = on_grid && in_block && o >= 0 && (npts_tab || o < ncells) &&
                     hst[hc] == EX_DEPOSIT;
    unsigned long long pend = __ballot(cnt);
    while (pend) {
        const int leader = __ffsll((unsigned long long)pend) - 1;
        const int val = __builtin_amdgcn_readlane(o, leader);
        const unsigned long long same = __ballot(cnt && o == val);
        if (lane == leader) atomicAdd(&hnp[hc], __popcll(same));
        pend &= ~same;
    }
}

// the FP64 search of the points k_owner deferred, one lane per point (the
// queue holds the points of each wave together, so lanes stay close)
__global__ __launch_bounds__(256) void k_owner_deferred(GridP gp, HashP h,
                                                        const unsigned *__restrict__ start,
                                                        const float4 *__restrict__ hp,
                                                        int *__restrict__ owner,
                                                        const unsigned *ctrs,
                                                        const unsigned *__restrict__ defq,
                                                        unsigned defq_cap,
                                                        const int *__restrict__ status,
                                                        int *__restrict__ npts, int ncells,
                                                        int *const *__restrict__ npts_tab,
                                                        const int *const *__restrict__ stex_tab)
{
    unsigned n = ctrs[C_DEFER];
    if (n > defq_cap) n = defq_cap;
    const unsigned nx = (unsigned)gp.bnum_idx[0], ny = (unsigned)gp.bnum_idx[1];
    for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const unsigned lidx = defq[i];
        const int lx = (int)(lidx % nx), ly = (int)((lidx / nx) % ny), lz = (int)(lidx / (nx * ny));
        const float px = idx2phys1(gp.bmin_idx[0] + lx, gp.step[0], gp.gmin[0]);
        const float py = idx2phys1(gp.bmin_idx[1] + ly, gp.step[1], gp.gmin[1]);
        const float pz = idx2phys1(gp.bmin_idx[2] + lz, gp.step[2], gp.gmin[2]);
        const int o = nearest_f64(px, py, pz, h, start, hp);
        owner[lidx] = o;
        if (o < 0) continue;
        const int hc = npts_tab ? (o & 0x7fffff) : o;
        const int *hst = npts_tab ? stex_tab[o >> 23] : status;
        int *hnp = npts_tab ? npts_tab[o >> 23] : npts;
        if ((npts_tab || o < ncells) && hst[hc] == EX_DEPOSIT) atomicAdd(&hnp[hc], 1);
    }
}

void launch_owner(const GridP &gp, const HashP &h, const unsigned *start, const float4 *hp,
                  int *owner, int prec, unsigned *ctrs, const int *status, int *npts, int ncells,
                  int *const *npts_tab, const int *const *stex_tab, unsigned *defq,
                  unsigned defq_cap, hipStream_t s)
{
    // one block: its counts start here; several: the caller clears all
    // blocks' counts before the first owner pass
    if (!npts_tab && ncells > 0) launch_fill_u32((unsigned *)npts, (size_t)ncells, 0u, s);
    const int tx = (gp.bnum_idx[0] + 3) / 4, ty = (gp.bnum_idx[1] + 3) / 4,
              tz = (gp.bnum_idx[2] + 3) / 4;
    const int ntiles = tx * ty * tz;
    if (ntiles <= 0) return;
    hipLaunchKernelGGL(k_owner, dim3((ntiles + 3) / 4), dim3(256), 0, s, gp, h, start, hp, owner,
                       prec, ctrs, tx, ty, ntiles, status, npts, ncells, npts_tab, stex_tab,
                       defq, prec == 0 ? defq_cap : 0u);
    if (prec == 0 && defq_cap > 0) {
        const unsigned nblk = defq_cap / 256 + 1;
        hipLaunchKernelGGL(k_owner_deferred, dim3(nblk < 2048 ? nblk : 2048), dim3(256), 0, s,
                           gp, h, start, hp, owner, ctrs, defq, defq_cap, status, npts, ncells,
                           npts_tab, stex_tab);
    }
}

// ---------------------------------------------------------------------------
// deposit status of each cell: its star is complete and every Voronoi vertex
// (circumcenter of a star tet) lies inside the data bounds +- 2 FLT_EPSILON
// (the reference's CellGridPts test, src/dense.cpp:1381-1392, on the exact
// bbox of the cell).  Both are "for all tets of the star" conditions, so they
// are decided tet by tet: one lane per tet flags its vertices
//   - incomplete: a hull face (neighbour -1) makes its three vertices hull
//     vertices (the star traversal of neighbor_edges would meet it);
//   - outside: its circumcenter leaves the bounds;
//   - defined: its circumcenter exists (flat tets of exactly cospherical
//     points have none and are skipped; a cell with none is degenerate).
// MIXED decides with the FP32 circumcenter and its error bound (cc.w) and
// recomputes in FP64 only when the bound straddles a threshold; FP64
// recomputes every circumcenter.  The flags are plain byte stores of 1.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int comp4(const int4 &q, int i)
{
    return i == 0 ? q.x : (i == 1 ? q.y : (i == 2 ? q.z : q.w));
}

__global__ void k_exact_flags(BlockIn in, GridP gp, int prec, unsigned char *__restrict__ flags,
                              unsigned *ctrs)
{
    const int nc = in.ncells;
    unsigned char *inc = flags, *out = flags + nc, *def = flags + 2 * (size_t)nc;
    float lo[3], hi[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const float e = (gp.dmax[i] - gp.dmin[i]) * 2.0f * FLT_EPSILON;
        lo[i] = gp.dmin[i] - e;
        hi[i] = gp.dmax[i] + e;
    }
    __shared__ unsigned s_fb;
    if (threadIdx.x == 0) s_fb = 0;
    __syncthreads();
    unsigned nfb = 0;   // FP64 re-checks of this lane
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < in.ntets;
         t += gridDim.x * blockDim.x) {
        const int4 vv = in.tets[2 * t];
        if (vv.x >= nc && vv.y >= nc && vv.z >= nc && vv.w >= nc) continue;   // ghosts only
        const int4 nb = in.tets[2 * t + 1];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            if (comp4(nb, i) >= 0) continue;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int u = comp4(vv, k);
                if (k != i && u < nc) inc[u] = 1;
            }
        }
        bool decided = false, is_out = false, is_def = false;
        if (prec == 0) {
            const float4 c = in.cc[t];
            const float err = c.w;   // NaN or inf: undecidable in float
            if (err >= 0.0f && err < 1e-3f) {
                bool surely_in = true, surely_out = false;
                const float q[3] = {c.x, c.y, c.z};
#pragma unroll
                for (int i = 0; i < 3; ++i) {
                    surely_in &= (q[i] - err > lo[i]) && (q[i] + err < hi[i]);
                    surely_out |= (q[i] + err < lo[i]) || (q[i] - err > hi[i]);
                }
                if (surely_in || surely_out) {
                    decided = true;
                    is_def = true;
                    is_out = surely_out;
                }
            }
            nfb += decided ? 0u : 1u;
        }
        if (!decided) {
            const float4 a = in.pts[vv.x], b = in.pts[vv.y], c = in.pts[vv.z], d = in.pts[vv.w];
            double ox, oy, oz;
            circumcenter_f64(a.x, a.y, a.z, b.x, b.y, b.z, c.x, c.y, c.z, d.x, d.y, d.z, ox, oy, oz);
            if (!(isnan(ox) || isnan(oy) || isnan(oz))) {
                is_def = true;
                is_out = ox < (double)lo[0] || oy < (double)lo[1] || oz < (double)lo[2] ||
                         ox > (double)hi[0] || oy > (double)hi[1] || oz > (double)hi[2];
            }
        }
        if (is_def) {
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int u = comp4(vv, k);
                if (u < nc) {
                    def[u] = 1;
                    if (is_out) out[u] = 1;
                }
            }
        }
    }
    // one global atomic per workgroup (and few workgroups): same-address
    // atomics serialize in L2, ~10 ns each, and one per wave outlasted the
    // kernel's own work
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) nfb += __shfl_xor(nfb, m);
    if ((threadIdx.x & 63) == 0 && nfb) atomicAdd(&s_fb, nfb);
    __syncthreads();
    if (threadIdx.x == 0 && s_fb) atomicAdd(&ctrs[C_FALLBACK], s_fb);
}

__global__ void k_exact_classify(const int *__restrict__ v2t, int nc,
                                 const unsigned char *__restrict__ flags, int *status)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nc) return;
    int st;
    if (v2t[v] < 0 || flags[v]) st = EX_INCOMPLETE;
    else if (!flags[2 * (size_t)nc + v]) st = EX_DEGENERATE;
    else if (flags[nc + v]) st = EX_OUTSIDE;
    else st = EX_DEPOSIT;
    status[v] = st;
}

void launch_exact_status(const BlockIn &in, const GridP &gp, int *status, int prec, unsigned *ctrs,
                         unsigned char *flags, hipStream_t s)
{
    if (in.ncells <= 0) return;
    launch_fill_u32((unsigned *)flags, ((size_t)in.ncells * 3 + 3) / 4, 0u, s);
    if (in.ntets > 0) {
        const int blocks = std::min(4096, (in.ntets + 255) / 256);
        hipLaunchKernelGGL(k_exact_flags, dim3(blocks), dim3(256), 0, s, in, gp, prec, flags, ctrs);
    }
    hipLaunchKernelGGL(k_exact_classify, dim3((in.ncells + 255) / 256), dim3(256), 0, s, in.v2t,
                       in.ncells, (const unsigned char *)flags, status);
}

// ---------------------------------------------------------------------------
// density, CIC of deposited cells without points (the point counts come from
// k_owner)
// ---------------------------------------------------------------------------

// With several blocks a point owned by another block's cell is that
// block's contribution, keyed like a received point (1 << 31 | rank << 24 |
// cell << 1), so that the fold orders it after this block's own cells.
__global__ void k_exact_density(GridP gp, const int *__restrict__ owner, const int *__restrict__ status,
                                const int *__restrict__ npts, int ncells, CellOut o, int prec,
                                int *const *__restrict__ npts_tab,
                                const int *const *__restrict__ stex_tab, int self)
{
    const size_t n = (size_t)gp.bnum_idx[0] * gp.bnum_idx[1] * gp.bnum_idx[2];
    for (size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x; g < n;
         g += (size_t)gridDim.x * blockDim.x) {
        const int ow = owner[g];
        if (ow < 0) continue;
        int N;
        unsigned key;
        if (npts_tab) {
            const int r = ow >> 23, cell = ow & 0x7fffff;
            if (stex_tab[r][cell] != EX_DEPOSIT) continue;
            N = npts_tab[r][cell];
            key = r == self ? cell_key(cell, 0)
                            : (0x80000000u | ((unsigned)r << 24) | cell_key(cell, 0));
        } else {
            if (ow >= ncells || status[ow] != EX_DEPOSIT) continue;
            N = npts[ow];
            key = cell_key(ow, 0);
        }
        const float m = prec == 0 ? gp.mass / (float)N : (float)((double)gp.mass / (double)N);
        // an owned point is in the block and its grid, and this is its first
        // contribution (every density pass runs before any CIC pass): the
        // first-contribution stores of push_point without its atomic
        o.count[g] = 1u;
        o.first[g] = make_uint2(key, __float_as_uint(m));
        o.density[g] = (float)((double)m / (double)gp.div);
    }
}

__global__ void k_exact_cic(BlockIn in, GridP gp, const int *__restrict__ status_ex,
                            CellOut o, const Router *rt, int self)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= in.ncells) return;
    const int s = status_ex[v];
    int out_status = s, out_n = 0;
    if (s == EX_DEPOSIT) {
        const int N = o.npts[v];
        if (N > 0) {
            out_status = 0;
            out_n = N;
        } else {
            const float4 p = in.pts[v];
            int mn[3];
            float m8[8];
            cic_weights(p.x, p.y, p.z, gp.mass, gp, mn, m8);
            for (int k = 0; k < 8; ++k)
                deposit_point(mn[0] + (k & 1), mn[1] + ((k >> 1) & 1), mn[2] + (k >> 2), m8[k],
                              cell_key(v, 0), gp, o, rt, self);
            out_status = 1;
            out_n = 8;
        }
    }
    o.status[v] = out_status;
    if (s != EX_DEPOSIT) o.npts[v] = 0;
    else o.npts[v] = out_n;
}

void launch_exact_tail(const BlockIn &in, const GridP &gp, const int *owner, const int *status_ex,
                       const CellOut &o, int prec, int *const *npts_tab,
                       const int *const *stex_tab, const Router *rt, int self, int parts,
                       hipStream_t s)
{
    const size_t n = (size_t)gp.bnum_idx[0] * gp.bnum_idx[1] * gp.bnum_idx[2];
    size_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    if (blocks == 0) blocks = 1;
    if (parts & 1)
        hipLaunchKernelGGL(k_exact_density, dim3((unsigned)blocks), dim3(256), 0, s, gp, owner,
                           status_ex, o.npts, in.ncells, o, prec, npts_tab, stex_tab, self);
    if ((parts & 2) && in.ncells > 0)
        hipLaunchKernelGGL(k_exact_cic, dim3((in.ncells + 255) / 256), dim3(256), 0, s, in, gp,
                           status_ex, o, rt, self);
}

}  // namespace t2
