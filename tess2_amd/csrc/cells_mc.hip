// cells_mc.hip -- the single-block cell kernel with several cells per wave.
//
// One 64-lane wavefront takes CPW consecutive cells (in the block's Morton
// order), G = 64 / CPW lanes per cell, for the geometry phase, where one
// cell per wave leaves most lanes idle:
//   1. star BFS (neighbor_edges' first-seen order, src/tet.cpp:228-270), the
//      same level-synchronous scheme as the one-cell kernel, per lane group:
//      a batch of G candidates finds or inserts its tets in the group's hash
//      table, the lowest candidate key wins each new tet, the winners take
//      the next slots in key order;
//   2. faces: the Delaunay neighbours in first-seen order, per group;
//   3. ring walks (fill_edge_link, src/tet.cpp:389-409), one face per lane:
//      Newell normal (src/dense.cpp:1131-1162), fan area with the double
//      square root (src/volume.cpp:32-46), distance to the neighbour;
//   4. volume (faces in neighbour order, src/volume.cpp:23-53) and the cell
//      bbox over the star's circumcenters (src/dense.cpp:701-714).
// The face planes then stay in LDS and the wave takes its cells one at a
// time, all 64 lanes on one cell, for the data-bounds filter, the PtInCell
// verdicts, the scanline search and the deposit (scan_deposit.hpp), whose
// control is wave-uniform.
//
// Each cell group owns BYTES of LDS: star records (verts), circumcenters,
// neighbour ids (tet ids as gathered, replaced by star slots as the BFS
// resolves them), a hash table of HG entries (keys, states) and two slot
// lists; after the geometry phase the group's face planes overlay its star
// records, and group 0's neighbour area holds the scan runs.  Cells beyond
// SG star tets or FG faces are deferred to the large-capacity pass.
#include "kernels.hpp"
#include "geom.hpp"
#include "deposit.hpp"
#include "scan_deposit.hpp"

namespace t2 {

namespace {

// group state codes (wave-uniform within a group)
enum { G_RUN = 0, G_NONE = 1, G_HULL = 2, G_DEFER = 3 };

__device__ __forceinline__ int rdl_i(int x, int l) { return __builtin_amdgcn_readlane(x, l); }
__device__ __forceinline__ float rdl_f(float x, int l)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), l));
}

__device__ __forceinline__ int comp4(const int4 &q, int i)
{
    return i == 0 ? q.x : (i == 1 ? q.y : (i == 2 ? q.z : q.w));
}

// Correctly rounded square root of a double holding a float value: the
// reciprocal square root refined by Goldschmidt's iteration with two Newton
// corrections, as the compiler expands sqrt(double), without the rescaling
// of inputs below 2^-767 that float values never need.
__device__ __forceinline__ double sqrt_of_float_mc(double x)
{
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y;
    double h = y * 0.5;
    const double r = __builtin_fma(-h, g, 0.5);
    g = __builtin_fma(g, r, g);
    double d = __builtin_fma(-g, g, x);
    h = __builtin_fma(h, r, h);
    g = __builtin_fma(d, h, g);
    d = __builtin_fma(-g, g, x);
    g = __builtin_fma(d, h, g);
    return __builtin_amdgcn_class(x, 0x260) ? x : g;
}

}  // namespace

template <int SG, int HG>
struct McLayout {
    static constexpr int V = 0;                    // int4 [SG] star tet vertices
    static constexpr int CC = 16 * SG;             // float4 [SG] circumcenters
    static constexpr int NB = 32 * SG;             // int4 [SG] neighbour tet ids -> slots
    static constexpr int HK = 48 * SG;             // int [HG] hash keys
    static constexpr int HST = 48 * SG + 4 * HG;   // int [HG] hash states
    static constexpr int TID = 48 * SG + 8 * HG;   // int [SG] slot -> tet id; face -> vertex; fvol
    static constexpr int AUX = 52 * SG + 8 * HG;   // int [SG] face -> first slot
    static constexpr int BYTES = 56 * SG + 8 * HG;
};

// One wave's CPW cells: v is the cell of this lane's group (-1 for none).
// Every lane of the wave calls it (it has wave barriers); smem holds CPW
// group regions of McLayout<SG, HG>::BYTES.  rt / self route off-block
// points to the other blocks of a multi-block run.
template <int CPW, int SG, int HG, int FG, int RMAX, bool WIDE, bool LINE>
__device__ __forceinline__ void cells_mc_body(const int v, const BlockIn &in, const GridP *gpp,
                                              const CellOut &o, const int want_density,
                                              const Router *rt, const int self,
                                              unsigned char *smem)
{
    constexpr int G = 64 / CPW;
    using L = McLayout<SG, HG>;
    static_assert(G * CPW == 64, "CPW divides 64");
    static_assert(FG <= SG && FG <= 64, "faces overlay the star records; one face per lane");
    static_assert(RMAX <= SG, "runs fit the neighbour area");
    constexpr int NFG = (FG + G - 1) / G;
    constexpr int LMAX = 4096;

    const int lane = (int)(threadIdx.x & 63);
    const int g = lane / G, gl = lane % G;
    const unsigned long long gbits =
        (G == 64) ? ~0ull : (((1ull << (G & 63)) - 1ull) << (g * G));
    const unsigned long long below = (1ull << lane) - 1ull;   // lanes under this one

    unsigned char *R = smem + g * L::BYTES;
    int4 *s_v = reinterpret_cast<int4 *>(R + L::V);
    float4 *s_cc = reinterpret_cast<float4 *>(R + L::CC);
    int *s_nb = reinterpret_cast<int *>(R + L::NB);
    int *hk = reinterpret_cast<int *>(R + L::HK);
    int *hst = reinterpret_cast<int *>(R + L::HST);
    int *s_tid = reinterpret_cast<int *>(R + L::TID);
    int *s_aux = reinterpret_cast<int *>(R + L::AUX);

    const int vt = v >= 0 ? in.v2t[v] : -1;
    int st = v < 0 ? G_NONE : (vt < 0 ? G_HULL : G_RUN);

    // ---- 1. star BFS, per group -------------------------------------------
    for (int i = gl; i < HG; i += G) {
        hk[i] = -1;
        hst[i] = 0x7fffffff;
    }
    __syncthreads();
    if (st == G_RUN && gl == 0) {
        s_v[0] = in.tets[2 * vt];
        reinterpret_cast<int4 *>(s_nb)[0] = in.tets[2 * vt + 1];
        s_cc[0] = in.cc[vt];
        hst[hash_find<HG>(hk, vt)] = -1;   // slot 0
        s_tid[0] = vt;
    }
    int S = 1, f0 = 0, f1 = 1;
    __syncthreads();
    // Each candidate lane loads its tet's record as soon as it knows the id,
    // before the table decides whether the tet is new: the loads are in
    // flight while the LDS hash resolves the batch, and the lanes that win
    // a slot write the record there -- no separate gather round per level.
    while (__ballot(st == G_RUN && f0 < f1)) {
        const bool act = st == G_RUN && f0 < f1;
        const int ncand = act ? (f1 - f0) * 4 : 0;
        for (int base = 0; __ballot(st == G_RUN && base < ncand); base += G) {
            const int key = base + gl;
            bool mine = st == G_RUN && key < ncand;
            int cand = -1, s = 0, i = 0;
            bool valid = false, h = false;
            int4 rv = make_int4(0, 0, 0, 0), rn = make_int4(0, 0, 0, 0);
            float4 rc = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (mine) {
                s = f0 + (key >> 2);
                i = key & 3;
                const int vi = reinterpret_cast<const int *>(s_v)[key + 4 * f0];
                const int ni = s_nb[key + 4 * f0];
                if (vi != v) {
                    if (ni < 0) h = true;
                    else {
                        valid = true;
                        cand = ni;
                        rv = in.tets[2 * ni];
                        rn = in.tets[2 * ni + 1];
                        rc = in.cc[ni];
                    }
                }
            }
            if (__ballot(h) & gbits) {   // a hull link: the cell is infinite
                st = G_HULL;
                mine = valid = false;
            }
            const int e = valid ? hash_find<HG>(hk, cand) : 0;
            if (__ballot(valid && e < 0) & gbits) {   // table full: star too large
                st = G_DEFER;
                mine = valid = false;
            }
            if (valid) atomicMin(&hst[e], key);
            const bool lead = valid && hst[e] == key;
            const unsigned long long Lb = __ballot(lead) & gbits;
            if (lead) {
                const int slot = S + __popcll(Lb & below);
                hst[e] = -(slot + 1);
                if (slot < SG) {
                    s_tid[slot] = cand;
                    s_v[slot] = rv;
                    reinterpret_cast<int4 *>(s_nb)[slot] = rn;
                    s_cc[slot] = rc;
                }
            }
            S += __popcll(Lb);
            if (mine) s_nb[s * 4 + i] = valid ? -hst[e] - 1 : -1;
            if (st == G_RUN && S > SG) st = G_DEFER;
        }
        __syncthreads();
        if (act && st == G_RUN) {
            f0 = f1;
            f1 = S;
        }
    }

    // ---- 2. faces: neighbours in first-seen order, per group --------------
    for (int i = gl; i < HG; i += G) {
        hk[i] = -1;
        hst[i] = 0x7fffffff;
    }
    int *s_fu = s_tid, *s_fs = s_aux;
    __syncthreads();
    int F = 0;
    {
        const int nent = st == G_RUN ? S * 4 : 0;
        for (int base = 0; __ballot(st == G_RUN && base < nent); base += G) {
            const int e = base + gl;
            int u = -1;
            bool valid = false;
            if (st == G_RUN && e < nent) {
                u = reinterpret_cast<const int *>(s_v)[e];
                valid = (u != v);
            }
            const int hx = valid ? hash_find<HG>(hk, u) : 0;
            if (__ballot(valid && hx < 0) & gbits) {
                st = G_DEFER;
                valid = false;
            }
            if (valid) atomicMin(&hst[hx], e);
            const bool lead = valid && hst[hx] == e;
            const unsigned long long Lb = __ballot(lead) & gbits;
            if (lead) {
                const int f = F + __popcll(Lb & below);
                hst[hx] = -(f + 1);
                if (f < FG) {
                    s_fu[f] = u;
                    s_fs[f] = e >> 2;
                }
            }
            F += __popcll(Lb);
            if (st == G_RUN && F > FG) st = G_DEFER;
        }
    }
    __syncthreads();
    int fu[NFG], fs[NFG];
#pragma unroll
    for (int j = 0; j < NFG; ++j) {
        const int f = j * G + gl;
        const bool have = st == G_RUN && f < F;
        fu[j] = have ? s_fu[f] : -1;
        fs[j] = have ? s_fs[f] : 0;
    }

    // ---- 3. ring walks: Newell normal, fan area, distance (one face per lane)
    float4 site = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (st == G_RUN) site = in.pts[v];
    float fnx[NFG], fny[NFG], fnz[NFG], ffx[NFG], ffy[NFG], ffz[NFG], fvol[NFG];
    bool ring_bad = false;
#pragma unroll
    for (int j = 0; j < NFG; ++j) {
        fnx[j] = fny[j] = fnz[j] = 0.0f;
        ffx[j] = ffy[j] = ffz[j] = 0.0f;
        fvol[j] = 0.0f;
        if (st == G_RUN && j * G + gl < F) {
            const int u = fu[j];
            const int us = fs[j];
            int4 tv = s_v[us];
            int wi = 0;
            for (int i = 3; i >= 0; --i) {   // circulate_start: first slot not v, u
                const int w = comp4(tv, i);
                if (w != v && w != u) wi = i;
            }
            const float4 c0 = s_cc[us];
            float cx = c0.x, cy = c0.y, cz = c0.z;
            float nx = 0.0f, ny = 0.0f, nz = 0.0f;
            float area = 0.0f;
            int t = us, Lr = 1;
            int behind = comp4(tv, wi);
            for (;;) {
                // circulate_next (src/tet.cpp:185-215): the next vertex is the
                // one of the tet's four that is not v, u or behind
                const int nvid = (int)((unsigned)tv.x + (unsigned)tv.y + (unsigned)tv.z +
                                       (unsigned)tv.w - (unsigned)v - (unsigned)u -
                                       (unsigned)behind);
                const int nt = s_nb[t * 4 + wi];
                if (nt == us || nt < 0) break;
                const int4 nv4 = s_v[nt];
                int nwi = -1;
                for (int i = 3; i >= 0; --i)
                    if (comp4(nv4, i) == nvid) nwi = i;
                const float4 c = s_cc[nt];
                // Newell term for (ring[L-1], ring[L])
                nx += (cy - c.y) * (cz + c.z);
                ny += (cz - c.z) * (cx + c.x);
                nz += (cx - c.x) * (cy + c.y);
                if (Lr >= 2) {   // fan triangle (ring[0], ring[L-1], ring[L])
                    const float abx = cx - c0.x, aby = cy - c0.y, abz = cz - c0.z;
                    const float acx = c.x - c0.x, acy = c.y - c0.y, acz = c.z - c0.z;
                    const float px = aby * acz - abz * acy;
                    const float py = abz * acx - abx * acz;
                    const float pz = abx * acy - aby * acx;
                    area = (float)((double)area +
                                   sqrt_of_float_mc((double)nrm3(px, py, pz)) / 2.0);
                }
                cx = c.x;
                cy = c.y;
                cz = c.z;
                t = nt;
                tv = nv4;
                wi = nwi;
                behind = nvid;
                Lr++;
                if (nwi < 0 || Lr > LMAX) {
                    ring_bad = (Lr > LMAX);
                    break;
                }
            }
            // closing Newell term (ring[L-1], ring[0])
            nx += (cy - c0.y) * (cz + c0.z);
            ny += (cz - c0.z) * (cx + c0.x);
            nz += (cx - c0.x) * (cy + c0.y);
            const float mag = sqrtf(nx * nx + ny * ny + nz * nz);
            nx = -(nx / mag);
            ny = -(ny / mag);
            nz = -(nz / mag);
            const float wx = c0.x - site.x, wy = c0.y - site.y, wz = c0.z - site.z;
            if (wx * nx + wy * ny + wz * nz < 0.0f) {
                nx = -nx;
                ny = -ny;
                nz = -nz;
            }
            fnx[j] = nx;
            fny[j] = ny;
            fnz[j] = nz;
            ffx[j] = c0.x;
            ffy[j] = c0.y;
            ffz[j] = c0.z;
            const float4 pu = in.pts[u];
            const float ddx = pu.x - site.x, ddy = pu.y - site.y, ddz = pu.z - site.z;
            const float dist = sqrtf(nrm3(ddx, ddy, ddz));
            fvol[j] = area * dist / 6.0f;
        }
    }
    if (__ballot(ring_bad) & gbits) st = G_DEFER;

    // ---- 4. cell bbox over the star's circumcenters, per group -------------
    float mnx = INFINITY, mny = INFINITY, mnz = INFINITY;
    float mxx = -INFINITY, mxy = -INFINITY, mxz = -INFINITY;
    if (st == G_RUN) {
        for (int q = gl; q < S; q += G) {
            const float4 c = s_cc[q];
            if (c.x < mnx) mnx = c.x;
            if (c.y < mny) mny = c.y;
            if (c.z < mnz) mnz = c.z;
            if (c.x > mxx) mxx = c.x;
            if (c.y > mxy) mxy = c.y;
            if (c.z > mxz) mxz = c.z;
        }
    }
#pragma unroll
    for (int off = G / 2; off >= 1; off >>= 1) {   // stays inside the group
        mnx = fminf(mnx, __shfl_xor(mnx, off));
        mny = fminf(mny, __shfl_xor(mny, off));
        mnz = fminf(mnz, __shfl_xor(mnz, off));
        mxx = fmaxf(mxx, __shfl_xor(mxx, off));
        mxy = fmaxf(mxy, __shfl_xor(mxy, off));
        mxz = fmaxf(mxz, __shfl_xor(mxz, off));
    }
    if (st == G_RUN) {
        // the first ring vertex of the first face seeds the reference's fold;
        // a NaN there survives every later comparison
        const float4 c00 = s_cc[s_fs[0]];
        if (isnan(c00.x)) { mnx = c00.x; mxx = c00.x; }
        if (isnan(c00.y)) { mny = c00.y; mxy = c00.y; }
        if (isnan(c00.z)) { mnz = c00.z; mxz = c00.z; }
    }
    __syncthreads();
    // volume: the faces in neighbour order (fvol staged over the face list)
    float *s_fvol = reinterpret_cast<float *>(s_tid);
#pragma unroll
    for (int j = 0; j < NFG; ++j) {
        const int f = j * G + gl;
        if (st == G_RUN && f < F) s_fvol[f] = fvol[j];
    }
    // face planes for the scan, over the star records (read for the last
    // time above)
    float4 *fa = reinterpret_cast<float4 *>(R + L::V);
    float2 *fb = reinterpret_cast<float2 *>(R + L::CC);
    __syncthreads();
#pragma unroll
    for (int j = 0; j < NFG; ++j) {
        const int f = j * G + gl;
        if (st == G_RUN && f < F) {
            fa[f] = make_float4(fnx[j], fny[j], fnz[j], ffx[j]);
            fb[f] = make_float2(ffy[j], ffz[j]);
        }
    }
    if (st == G_RUN && gl == 0 && o.vol) {
        float vol = 0.0f;
        for (int q = 0; q < F; ++q) vol += s_fvol[q];
        o.vol[v] = vol;
    }
    __syncthreads();

    // ---- 5. the cells one at a time, the whole wave on each ----------------
    int4 *runs = reinterpret_cast<int4 *>(smem + L::NB);   // group 0's neighbour area
    for (int q = 0; q < CPW; ++q) {
        const int ql = q * G;
        const int vq = rdl_i(v, ql);
        const int sq = rdl_i(st, ql);
        if (sq == G_NONE) continue;
        if (sq == G_HULL) {
            if (lane == 0) {
                o.status[vq] = ST_INCOMPLETE;
                o.npts[vq] = 0;
                if (o.vol) o.vol[vq] = -1.0f;
            }
            continue;
        }
        if (sq == G_DEFER) {
            if (lane == 0) {
                const unsigned kk = atomicAdd(&o.ctrs[C_SLOW], 1u);
                if (kk < o.slow_cap) o.slow[kk] = vq;
                else atomicOr(&o.ctrs[C_ERR], 1u);   // list full: rerun with more room
            }
            continue;
        }
        if (!want_density) {
            if (lane == 0) {
                o.status[vq] = ST_SCAN;
                o.npts[vq] = 0;
            }
            continue;
        }
        const int Fq = rdl_i(F, ql);
        unsigned char *Rq = smem + q * L::BYTES;
        const float4 *faq = reinterpret_cast<const float4 *>(Rq + L::V);
        const float2 *fbq = reinterpret_cast<const float2 *>(Rq + L::CC);
        float nx1[1] = {0.0f}, ny1[1] = {0.0f}, nz1[1] = {0.0f};
        float fx1[1] = {0.0f}, fy1[1] = {0.0f}, fz1[1] = {0.0f};
        if (lane < Fq) {
            const float4 a = faq[lane];
            const float2 b = fbq[lane];
            nx1[0] = a.x;
            ny1[0] = a.y;
            nz1[0] = a.z;
            fx1[0] = a.w;
            fy1[0] = b.x;
            fz1[0] = b.y;
        }
        const float4 siteq = in.pts[vq];
        scan_deposit<1, RMAX, WIDE, LINE>(vq, siteq, Fq, nx1, ny1, nz1, fx1, fy1, fz1,
                                     rdl_f(mnx, ql), rdl_f(mny, ql), rdl_f(mnz, ql),
                                     rdl_f(mxx, ql), rdl_f(mxy, ql), rdl_f(mxz, ql), *gpp, o,
                                     rt, self, const_cast<float4 *>(faq),
                                     const_cast<float2 *>(fbq), true, runs);
        __syncthreads();
    }
}

template <int CPW, int SG, int HG, int FG, int RMAX, int MINW, bool WIDE>
__global__ __launch_bounds__(64, MINW) void k_cells_mc(BlockIn in,
                                                       const GridP *__restrict__ gp, CellOut o,
                                                       int want_density)
{
    __shared__ __align__(16) unsigned char smem[CPW * McLayout<SG, HG>::BYTES];
    const int nitems = (in.ncells + CPW - 1) / CPW;
    const int item = xcd_item((int)blockIdx.x, nitems);
    if (item < 0) return;
    const int k = item * CPW + (int)(threadIdx.x & 63) / (64 / CPW);
    const int v = k < in.ncells ? (in.order ? in.order[k] : k) : -1;
    cells_mc_body<CPW, SG, HG, FG, RMAX, WIDE, false>(v, in, gp, o, want_density, nullptr, 0,
                                                      smem);
}

// All blocks of a run in one persistent launch (as k_cells_ref_batched):
// waves take chunks of consecutive cells of one block from their XCD's
// queue, then from the others, and run CPW cells of the chunk at a time.
template <int CPW, int SG, int HG, int FG, int RMAX, int MINW, bool WIDE>
__global__ __launch_bounds__(64, MINW) void k_cells_mc_batched(
    const BlockIn *__restrict__ ins, const GridP *__restrict__ gps,
    const CellOut *__restrict__ outs, const Router *__restrict__ rt,
    const Chunk *__restrict__ chunks, XcdQueues xq, unsigned *queue, int want_density)
{
    __shared__ __align__(16) unsigned char smem[CPW * McLayout<SG, HG>::BYTES];
    const int lane = (int)(threadIdx.x & 63);
    const int g = lane / (64 / CPW);
    const int x0 = (int)(blockIdx.x & 7);
    for (int q = 0; q < 8;) {
        const int x = (x0 + q) & 7;
        int c = 0;
        if (lane == 0) c = (int)atomicAdd(&queue[x], 1u);
        c = xq.start[x] + __builtin_amdgcn_readfirstlane(c);
        if (c >= xq.start[x + 1]) {   // range x is done: help the next one
            ++q;
            continue;
        }
        const Chunk ch = chunks[c];
        const BlockIn in = ins[ch.block];
        const CellOut o = outs[ch.block];
        for (int j = 0; j < ch.count; j += CPW) {
            const int kk = ch.start + (j + g) * ch.stride;
            const int v = j + g < ch.count ? (in.order ? in.order[kk] : kk) : -1;
            cells_mc_body<CPW, SG, HG, FG, RMAX, WIDE, true>(v, in, gps + ch.block, o,
                                                             want_density, rt, ch.block, smem);
            __syncthreads();
        }
    }
}

// All blocks of a run in one launch without a work queue: workgroup b runs
// item xcd_item(b) of the item list (CPW consecutive cells of one block, the
// blocks in staging order: each XCD gets one contiguous part of the list),
// with the block's descriptors read per workgroup.  The dispatcher balances
// the waves within an XCD, not across them: this beats the persistent
// kernels where the blocks cost alike (8 uniform blocks: 5.6 vs 7.3 ms) and
// loses where they do not (64 Plummer blocks: 24.1 vs 19.7 ms); dealing the
// items to the XCDs in tiles instead loses L2 locality (35 ms there).
template <int CPW, int SG, int HG, int FG, int RMAX, int MINW, bool WIDE>
__global__ __launch_bounds__(64, MINW) void k_cells_mc_items(
    const BlockIn *__restrict__ ins, const GridP *__restrict__ gps,
    const CellOut *__restrict__ outs, const Router *__restrict__ rt,
    const Chunk *__restrict__ items, XcdQueues xr, int want_density)
{
    __shared__ __align__(16) unsigned char smem[CPW * McLayout<SG, HG>::BYTES];
    const int x = (int)(blockIdx.x & 7);
    int s0 = xr.start[0], s1 = xr.start[1];
#pragma unroll
    for (int i = 1; i < 8; ++i)
        if (x == i) {
            s0 = xr.start[i];
            s1 = xr.start[i + 1];
        }
    const int item = s0 + (int)(blockIdx.x >> 3);
    if (item >= s1) return;
    const Chunk ch = items[item];
    const BlockIn in = ins[ch.block];
    const CellOut o = outs[ch.block];
    const int g = (int)(threadIdx.x & 63) / (64 / CPW);
    const int kk = ch.start + g * ch.stride;
    const int v = g < ch.count ? (in.order ? in.order[kk] : kk) : -1;
    cells_mc_body<CPW, SG, HG, FG, RMAX, WIDE, true>(v, in, gps + ch.block, o, want_density, rt,
                                                     ch.block, smem);
}

void launch_cells_mc_items(const BlockIn *ins, const GridP *gps, const CellOut *outs,
                           const Router *rt, const Chunk *items, const XcdQueues &xr,
                           int want_density, hipStream_t s)
{
    int most = 0;
    for (int x = 0; x < 8; ++x) most = std::max(most, xr.start[x + 1] - xr.start[x]);
    if (most <= 0) return;
    // 6 waves per SIMD (80 VGPRs, 2 spilled): 5.6 ms on 8 uniform blocks
    // against 6.4 at 5 waves (no spills) and 6.3 at 4
    hipLaunchKernelGGL((k_cells_mc_items<2, 48, 64, 48, 48, 6, true>), dim3(8 * most),
                       dim3(64), 0, s, ins, gps, outs, rt, items, xr, want_density);
}

template <int CPW, int SG, int HG, int FG, int RMAX, int MINW = 1, bool WIDE = false>
static void launch_mc(const BlockIn &in, const GridP *gp, const CellOut &o, int want_density,
                      hipStream_t s)
{
    const int nitems = (in.ncells + CPW - 1) / CPW;
    hipLaunchKernelGGL((k_cells_mc<CPW, SG, HG, FG, RMAX, MINW, WIDE>), dim3(xcd_grid(nitems)),
                       dim3(64), 0, s, in, gp, o, want_density);
}

void launch_cells_mc_batched(const BlockIn *ins, const GridP *gps, const CellOut *outs,
                             const Router *rt, const Chunk *chunks, const XcdQueues &xq,
                             unsigned *queue, int waves, int want_density, hipStream_t s)
{
    launch_fill_u32(queue, 8, 0u, s);
    if (xq.start[8] > 0)
        hipLaunchKernelGGL((k_cells_mc_batched<2, 48, 64, 48, 48, 6, true>), dim3(waves), dim3(64),
                           0, s, ins, gps, outs, rt, chunks, xq, queue, want_density);
}

int launch_cells_mc(const BlockIn &in, const GridP *gp, const CellOut &o, int want_density,
                    int variant, hipStream_t s)
{
    if (in.ncells <= 0) return 0;
    // two cells per wave, 48-tet stars, 64-entry tables: 6 waves per SIMD
    // (76 VGPRs, 6.4 KB LDS); measured best among 1, 2 and 4 cells per wave
    // and star capacities 40..64.  Variant 2 adds the multi-word verdict
    // masks for boxes of up to 512 points.
    switch (variant) {
    case 1: launch_mc<2, 48, 64, 48, 48, 1, false>(in, gp, o, want_density, s); return 0;
    case 2: launch_mc<2, 48, 64, 48, 48, 1, true>(in, gp, o, want_density, s); return 0;
    default: return -1;
    }
}

}  // namespace t2
