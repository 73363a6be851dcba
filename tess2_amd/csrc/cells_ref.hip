// cells_ref.hip -- gfx950 kernels of the reference-exact density path.
//
// One 64-lane wavefront handles one Voronoi cell (one original particle):
//   1. star: level-synchronous BFS over the tets around the site, starting at
//      vert_to_tet[v], that visits tets in exactly the first-seen order of
//      neighbor_edges() (src/tet.cpp:228-270); each BFS level is one round of
//      parallel gathers, duplicates are resolved through a tet-id -> slot
//      hash table in LDS.  Star tet records and their circumcenters are
//      staged in LDS.
//   2. faces: the Delaunay neighbours u in first-seen order with their first
//      tet ut (one face per lane); each lane walks its edge ring in LDS
//      (fill_edge_link, src/tet.cpp:389-409) and accumulates, in ring order,
//      the Newell normal (src/dense.cpp:1131-1162) and the fan area
//      (src/volume.cpp:32-46).
//   3. volume = sum of area*|u-v|/6 over faces in neighbour order
//      (src/volume.cpp:23-53).
//   4. cell bbox over the star's circumcenters (the ring vertices), data-
//      bounds filter, bbox -> grid range (src/dense.cpp:701-714,1381-1410).
//   5. scanline search (CellInteriorGridPts, src/dense.cpp:1475-1701) as a
//      wave-uniform state machine.  PtInCell (src/dense.cpp:1172-1203: inside
//      iff no two significant faces disagree in sign) is evaluated for the
//      whole cell box at once when it is small, one lane per point, into
//      bit masks whose lines the walks scan with bit operations; larger
//      boxes probe point by point, all faces at once (one per lane) with
//      two ballots per verdict.
//   6. deposit: mass/N per found point (or CIC of the site when N = 0).
//      The first contribution to a grid point writes the density directly;
//      later ones go to a shared-point list that an ordered pass folds in
//      ascending cell id, replaying density += mass/div (float += double,
//      src/dense.cpp:290) in the reference's order.
#include "kernels.hpp"
#include "geom.hpp"
#include "deposit.hpp"
#include "scan_bits.hpp"
#include "scan_deposit.hpp"

namespace t2 {

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & 63); }

__device__ __forceinline__ int rdl(int x, int l) { return __builtin_amdgcn_readlane(x, l); }
__device__ __forceinline__ float rdlf(float x, int l)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), l));
}

// Correctly rounded square root of a double that holds a float value (so
// never below 2^-149 unless zero): Goldschmidt's iteration from the hardware
// reciprocal square root with two Newton corrections of the root, as the
// compiler expands sqrt(double), without the rescaling of inputs below
// 2^-767 that such values never need.  Zero and +inf return themselves.
__device__ __forceinline__ double sqrt_of_float(double x)
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

__device__ __forceinline__ int comp(const int4 &q, int i)
{
    return i == 0 ? q.x : (i == 1 ? q.y : (i == 2 ? q.z : q.w));
}

// ---------------------------------------------------------------------------
// the per-cell wavefront kernel
//   NS   : star capacity in units of 64 tets
//   NF   : face capacity in units of 64 faces
//   RMAX : scan runs kept in LDS (a cell with more scans twice, see step 5)
// ---------------------------------------------------------------------------
// LDS ints of the star / face tables: HS keys, HS states, 2 x SMAX slots
template <int NS> constexpr int hash_entries() { return 2 * 64 * NS; }
template <int NS> constexpr int hash_ints() { return 2 * hash_entries<NS>() + 2 * 64 * NS; }

template <int NS, int NF, int RMAX, bool WIDE = false, bool SPEC = false>
__device__ void cell_ref_body(const int v, const BlockIn &in, const GridP *gpp,
                              const CellOut &o, const int want_density, const bool can_defer,
                              int4 *s_v, int *s_ns, float4 *s_cc, int4 *s_u, int *s_h,
                              const Router *rt = nullptr, const int self = 0)
{
    constexpr int SMAX = 64 * NS;
    constexpr int FMAX = 64 * NF;
    constexpr int HS = hash_entries<NS>();
    static_assert(FMAX <= SMAX, "face lists reuse the slot arrays");
    constexpr int LMAX = 4096;
    const int lane = lane_id();

    const int vt = in.v2t[v];
    if (vt < 0) {
        if (lane == 0) {
            o.status[v] = ST_INCOMPLETE;
            o.npts[v] = 0;
            if (o.vol) o.vol[v] = -1.0f;
        }
        return;
    }

    // ---- 1. star BFS -------------------------------------------------------
    // Tet ids map to star slots through an open-addressing table in LDS.  A
    // batch of up to 64 candidates resolves at once: every lane finds or
    // inserts its tet, an atomicMin leaves the lowest candidate key on each
    // tet first seen in this batch, and those leaders take the next slots in
    // key order -- the order in which neighbor_edges() first meets them.
    int *hk = s_h, *hst = s_h + HS, *s_tid = s_h + 2 * HS, *s_aux = s_tid + SMAX;
    for (int i = lane; i < HS; i += 64) {
        hk[i] = -1;
        hst[i] = 0x7fffffff;
    }
    __syncthreads();
    if (lane == 0) {
        s_v[0] = in.tets[2 * vt];
        s_u[0] = in.tets[2 * vt + 1];
        s_cc[0] = in.cc[vt];
        hst[hash_find<HS>(hk, vt)] = -1;   // slot 0
        s_tid[0] = vt;
    }
    int S = 1, f0 = 0, f1 = 1;
    bool hull = false, over = false;
    __syncthreads();
    while (f0 < f1) {
        const int ncand = (f1 - f0) * 4;
        for (int base = 0; base < ncand; base += 64) {
            const int key = base + lane;
            int cand = -1, s = 0, i = 0;
            bool valid = false, h = false;
            int4 rv = make_int4(0, 0, 0, 0), rn = make_int4(0, 0, 0, 0);
            float4 rc = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (key < ncand) {
                s = f0 + (key >> 2);
                i = key & 3;
                const int vi = reinterpret_cast<const int *>(s_v)[key + 4 * f0];
                const int ni = reinterpret_cast<const int *>(s_u)[key + 4 * f0];
                if (vi != v) {
                    if (ni < 0) h = true;
                    else {
                        valid = true;
                        cand = ni;
                        if (SPEC) {   // the record's loads overlap the table work below
                            rv = in.tets[2 * ni];
                            rn = in.tets[2 * ni + 1];
                            rc = in.cc[ni];
                        }
                    }
                }
            }
            if (__ballot(h)) {
                hull = true;
                break;
            }
            const int e = valid ? hash_find<HS>(hk, cand) : 0;
            if (__ballot(valid && e < 0)) {   // table full: the star is too large
                over = true;
                break;
            }
            if (valid) atomicMin(&hst[e], key);
            const bool lead = valid && hst[e] == key;
            const unsigned long long L = __ballot(lead);
            if (lead) {
                const int slot = S + __popcll(L & ((1ull << lane) - 1ull));
                hst[e] = -(slot + 1);
                if (slot < SMAX) {
                    s_tid[slot] = cand;
                    if (SPEC) {
                        s_v[slot] = rv;
                        s_u[slot] = rn;
                        s_cc[slot] = rc;
                    }
                }
            }
            S += __popcll(L);
            if (key < ncand) s_ns[s * 4 + i] = valid ? -hst[e] - 1 : -1;
            if (S > SMAX) {
                over = true;
                break;
            }
        }
        if (hull || over) break;
        __syncthreads();
        // gather the records of the new slots [f1, S) (already written by
        // the winning lanes when SPEC)
        if (!SPEC) {
            for (int b = f1; b < S; b += 64) {
                const int s = b + lane;
                if (s < S) {
                    const int id = s_tid[s];
                    s_v[s] = in.tets[2 * id];
                    s_u[s] = in.tets[2 * id + 1];
                    s_cc[s] = in.cc[id];
                }
            }
        }
        __syncthreads();
        f0 = f1;
        f1 = S;
    }
    if (hull) {
        if (lane == 0) {
            o.status[v] = ST_INCOMPLETE;
            o.npts[v] = 0;
            if (o.vol) o.vol[v] = -1.0f;
        }
        return;
    }
    if (over) goto defer;

    {
        // ---- 2. faces: neighbours in first-seen order -----------------------
        // same table scheme keyed by vertex id; the entry key e = 4*slot + i
        // orders first sightings, and a face's first tet is e >> 2
        for (int i = lane; i < HS; i += 64) {
            hk[i] = -1;
            hst[i] = 0x7fffffff;
        }
        int *s_fu = s_tid, *s_fs = s_aux;
        __syncthreads();
        int F = 0;
        const int nent = S * 4;
        for (int base = 0; base < nent; base += 64) {
            const int e = base + lane;
            int u = -1;
            bool valid = false;
            if (e < nent) {
                u = reinterpret_cast<const int *>(s_v)[e];
                valid = (u != v);
            }
            const int hx = valid ? hash_find<HS>(hk, u) : 0;
            if (__ballot(valid && hx < 0)) {
                F = FMAX + 1;
                break;
            }
            if (valid) atomicMin(&hst[hx], e);
            const bool lead = valid && hst[hx] == e;
            const unsigned long long L = __ballot(lead);
            if (lead) {
                const int f = F + __popcll(L & ((1ull << lane) - 1ull));
                hst[hx] = -(f + 1);
                if (f < FMAX) {
                    s_fu[f] = u;
                    s_fs[f] = e >> 2;
                }
            }
            F += __popcll(L);
            if (F > FMAX) break;
        }
        if (F > FMAX) goto defer;
        __syncthreads();
        int fu[NF], fs[NF];
#pragma unroll
        for (int j = 0; j < NF; ++j) {
            const int f = j * 64 + lane;
            fu[j] = f < F ? s_fu[f] : -1;
            fs[j] = f < F ? s_fs[f] : 0;
        }

        // ---- ring walks: Newell normal, fan area, bbox (one face per lane) --
        const float4 site = in.pts[v];
        float fnx[NF], fny[NF], fnz[NF], ffx[NF], ffy[NF], ffz[NF], fvol[NF];
        float mnx = INFINITY, mny = INFINITY, mnz = INFINITY;
        float mxx = -INFINITY, mxy = -INFINITY, mxz = -INFINITY;
        bool ring_bad = false;
#pragma unroll
        for (int j = 0; j < NF; ++j) {
            fnx[j] = fny[j] = fnz[j] = 0.0f;
            ffx[j] = ffy[j] = ffz[j] = 0.0f;
            fvol[j] = 0.0f;
            if (j * 64 + lane < F) {
                const int u = fu[j];
                const int us = fs[j];
                int4 tv = s_v[us];
                int wi = 0;
                for (int i = 3; i >= 0; --i) {   // circulate_start: first slot not v,u
                    const int w = comp(tv, i);
                    if (w != v && w != u) wi = i;
                }
                const float4 c0 = s_cc[us];
                float cx = c0.x, cy = c0.y, cz = c0.z;
                float nx = 0.0f, ny = 0.0f, nz = 0.0f;
                float area = 0.0f;
                int t = us, L = 1;
                // the vertex at wi: the ring walk crosses the face opposite it
                int behind = comp(tv, wi);
                for (;;) {
                    // circulate_next (src/tet.cpp:185-215): the next vertex is
                    // the one of the tet's four that is not v, u or behind
                    const int nvid = (int)((unsigned)tv.x + (unsigned)tv.y + (unsigned)tv.z +
                                           (unsigned)tv.w - (unsigned)v - (unsigned)u -
                                           (unsigned)behind);
                    const int nt = s_ns[t * 4 + wi];
                    if (nt == us || nt < 0) break;
                    const int4 nv4 = s_v[nt];
                    int nwi = -1;
                    for (int i = 3; i >= 0; --i)
                        if (comp(nv4, i) == nvid) nwi = i;
                    const float4 c = s_cc[nt];
                    // Newell term for (ring[L-1], ring[L])
                    nx += (cy - c.y) * (cz + c.z);
                    ny += (cz - c.z) * (cx + c.x);
                    nz += (cx - c.x) * (cy + c.y);
                    if (L >= 2) {   // fan triangle (ring[0], ring[L-1], ring[L])
                        const float abx = cx - c0.x, aby = cy - c0.y, abz = cz - c0.z;
                        const float acx = c.x - c0.x, acy = c.y - c0.y, acz = c.z - c0.z;
                        const float px = aby * acz - abz * acy;
                        const float py = abz * acx - abx * acz;
                        const float pz = abx * acy - aby * acx;
                        area = (float)((double)area + sqrt_of_float((double)nrm3(px, py, pz)) / 2.0);
                    }
                    cx = c.x;
                    cy = c.y;
                    cz = c.z;
                    t = nt;
                    tv = nv4;
                    wi = nwi;
                    behind = nvid;
                    L++;
                    if (nwi < 0 || L > LMAX) {
                        ring_bad = (L > LMAX);
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
        if (__ballot(ring_bad)) goto defer;
        // cell bbox (src/dense.cpp:701-714): every ring vertex is the
        // circumcenter of a star tet and every star tet lies on the rings of
        // its three edges from the site, so the bbox of the rings is the bbox
        // of the star's circumcenters -- one lane per tet here
        for (int k = lane; k < S; k += 64) {
            const float4 c = s_cc[k];
            if (c.x < mnx) mnx = c.x;
            if (c.y < mny) mny = c.y;
            if (c.z < mnz) mnz = c.z;
            if (c.x > mxx) mxx = c.x;
            if (c.y > mxy) mxy = c.y;
            if (c.z > mxz) mxz = c.z;
        }

        // ---- 3. volume: faces summed in neighbour order ----------------------
        if (o.vol) {
            float vol = 0.0f;
            for (int k = 0; k < F; ++k) {
                float x = 0.0f;
#pragma unroll
                for (int j = 0; j < NF; ++j)
                    if ((k >> 6) == j) x = rdlf(fvol[j], k & 63);
                vol += x;
            }
            if (lane == 0) o.vol[v] = vol;
        }
        if (!want_density) {
            if (lane == 0) {
                o.status[v] = ST_SCAN;
                o.npts[v] = 0;
            }
            return;
        }

        // ---- 4. cell bbox, data-bounds filter (src/dense.cpp:701-714,1381-1392)
        for (int off = 32; off >= 1; off >>= 1) {
            mnx = fminf(mnx, __shfl_xor(mnx, off));
            mny = fminf(mny, __shfl_xor(mny, off));
            mnz = fminf(mnz, __shfl_xor(mnz, off));
            mxx = fmaxf(mxx, __shfl_xor(mxx, off));
            mxy = fmaxf(mxy, __shfl_xor(mxy, off));
            mxz = fmaxf(mxz, __shfl_xor(mxz, off));
        }
        {
            // the first ring vertex of the first face seeds the reference's
            // fold; a NaN there survives every later comparison
            const float4 c00 = s_cc[rdl(fs[0], 0)];
            if (isnan(c00.x)) { mnx = c00.x; mxx = c00.x; }
            if (isnan(c00.y)) { mny = c00.y; mxy = c00.y; }
            if (isnan(c00.z)) { mnz = c00.z; mxz = c00.z; }
        }
        scan_deposit<NF, RMAX, WIDE, WIDE>(v, site, F, fnx, fny, fnz, ffx, ffy, ffz, mnx, mny, mnz,
                                     mxx, mxy, mxz, *gpp, o, rt, self, s_cc,
                                     reinterpret_cast<float2 *>(s_v), false, s_u);
        return;
    }

defer:
    if (lane == 0) {
        if (can_defer) {
            const unsigned k = atomicAdd(&o.ctrs[C_SLOW], 1u);
            if (k < o.slow_cap) o.slow[k] = v;
            else atomicOr(&o.ctrs[C_ERR], 1u);   // list full: rerun with more room
        } else {
            atomicOr(&o.ctrs[C_ERR], 4u);        // beyond the large-capacity pass
            o.status[v] = -1;
            o.npts[v] = 0;
        }
    }
}

template <int NS, int NF, int RMAX>
__global__ __launch_bounds__(64) void k_cells_ref(BlockIn in, GridP gp, CellOut o,
                                                  int want_density)
{
    __shared__ int4 s_v[64 * NS];
    __shared__ int s_ns[64 * NS * 4];
    __shared__ float4 s_cc[64 * NS];
    __shared__ int4 s_u[(64 * NS > RMAX) ? 64 * NS : RMAX];
    __shared__ int s_h[hash_ints<NS>()];
    const int k = xcd_item((int)blockIdx.x, in.ncells);
    if (k < 0) return;
    const int v = in.order ? in.order[k] : k;
    cell_ref_body<NS, NF, RMAX>(v, in, &gp, o, want_density, true, s_v, s_ns, s_cc, s_u, s_h);
}

// large-capacity pass over the cells the fast kernel deferred; the grid is
// fixed and each workgroup strides over the deferred list
template <int NS, int NF, int RMAX>
__global__ __launch_bounds__(64) void k_cells_ref_slow(BlockIn in, GridP gp, CellOut o,
                                                       int want_density)
{
    __shared__ int4 s_v[64 * NS];
    __shared__ int s_ns[64 * NS * 4];
    __shared__ float4 s_cc[64 * NS];
    __shared__ int4 s_u[(64 * NS > RMAX) ? 64 * NS : RMAX];
    __shared__ int s_h[hash_ints<NS>()];
    unsigned n = o.ctrs[C_SLOW];
    if (n > o.slow_cap) n = o.slow_cap;
    for (unsigned k = blockIdx.x; k < n; k += gridDim.x) {
        const int v = o.slow[k];
        cell_ref_body<NS, NF, RMAX>(v, in, &gp, o, want_density, false, s_v, s_ns, s_cc, s_u,
                                    s_h);
        __syncthreads();
    }
}

// All blocks of a run in one persistent launch: every wave takes chunks of
// cells (one block each, consecutive in the block's Morton order) from the
// queue of its XCD's range of the chunk list, then from the other ranges,
// until all are empty.  Chunk sizes by block size class (tess2_autotune).  At least 6
// waves per SIMD: the per-block descriptors read from memory otherwise push
// the register count down to 4 waves (30 vs 34.5 ms on 64 Plummer blocks).
template <int NS, int NF, int RMAX, bool SPEC>
__global__ __launch_bounds__(64, 6) void k_cells_ref_batched(const BlockIn *__restrict__ ins,
                                                          const GridP *__restrict__ gps,
                                                          const CellOut *__restrict__ outs,
                                                          const Router *__restrict__ rt,
                                                          const Chunk *__restrict__ chunks,
                                                          XcdQueues xq, unsigned *queue,
                                                          int want_density)
{
    __shared__ int4 s_v[64 * NS];
    __shared__ int s_ns[64 * NS * 4];
    __shared__ float4 s_cc[64 * NS];
    __shared__ int4 s_u[(64 * NS > RMAX) ? 64 * NS : RMAX];
    __shared__ int s_h[hash_ints<NS>()];
    const int lane = lane_id();
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
        // read-only descriptors (scalar loads; rt->gps/outs alias them for routing)
        const BlockIn in = ins[ch.block];
        // a register copy: read from memory inside the cell body this
        // kernel measured 6 % slower (the two-cell kernels gain from it)
        const GridP gp = gps[ch.block];
        const CellOut o = outs[ch.block];
        for (int j = 0; j < ch.count; ++j) {
            const int k = ch.start + j * ch.stride;
            const int v = in.order ? in.order[k] : k;
            cell_ref_body<NS, NF, RMAX, true, SPEC>(v, in, &gp, o, want_density, true, s_v, s_ns,
                                                    s_cc, s_u,
                                        s_h, rt, ch.block);
            __syncthreads();
        }
    }
}

template <int NS, int NF, int RMAX>
__global__ __launch_bounds__(64) void k_cells_ref_slow_batched(const BlockIn *__restrict__ ins,
                                                               const Router *__restrict__ rt,
                                                               int want_density)
{
    __shared__ int4 s_v[64 * NS];
    __shared__ int s_ns[64 * NS * 4];
    __shared__ float4 s_cc[64 * NS];
    __shared__ int4 s_u[(64 * NS > RMAX) ? 64 * NS : RMAX];
    __shared__ int s_h[hash_ints<NS>()];
    // the lists of up to 64 blocks at a time, concatenated and dealt to the
    // workgroups in turn (one block per lane: list lengths and their
    // inclusive prefix sum), so that no workgroup takes one cell of every
    // block while others idle
    const int lane = lane_id();
    const int nb = rt->nblocks;
    for (int b0 = 0; b0 < nb; b0 += 64) {
        unsigned n = 0;
        if (b0 + lane < nb) {
            const CellOut &o = rt->outs[b0 + lane];
            n = min(o.ctrs[C_SLOW], o.slow_cap);
        }
        unsigned incl = n;
        for (int d = 1; d < 64; d <<= 1) {
            const unsigned t = __shfl_up(incl, d);
            if (lane >= d) incl += t;
        }
        const unsigned total = __shfl(incl, 63);
        for (unsigned g = blockIdx.x; g < total; g += gridDim.x) {
            const int lb = __ffsll((unsigned long long)__ballot(incl > g)) - 1;
            const unsigned first = __shfl(incl - n, lb);
            const int b = b0 + lb;
            const CellOut &o = rt->outs[b];
            cell_ref_body<NS, NF, RMAX, true>(o.slow[g - first], ins[b], rt->gps + b, o,
                                              want_density, false, s_v, s_ns, s_cc, s_u, s_h,
                                              rt, b);
            __syncthreads();
        }
    }
}

void launch_cells_ref_batched(const BlockIn *ins, const GridP *gps, const CellOut *outs,
                              const Router *rt, const Chunk *chunks, const XcdQueues &xq,
                              unsigned *queue, int waves, int want_density, hipStream_t s,
                              bool spec)
{
    launch_fill_u32(queue, 8, 0u, s);
    if (xq.start[8] > 0) {
        if (spec)
            hipLaunchKernelGGL((k_cells_ref_batched<1, 1, 64, true>), dim3(waves), dim3(64), 0, s,
                               ins, gps, outs, rt, chunks, xq, queue, want_density);
        else
            hipLaunchKernelGGL((k_cells_ref_batched<1, 1, 64, false>), dim3(waves), dim3(64), 0, s,
                               ins, gps, outs, rt, chunks, xq, queue, want_density);
    }
    launch_cells_ref_slow_batched(ins, rt, want_density, s);
}

void launch_cells_ref_slow_batched(const BlockIn *ins, const Router *rt, int want_density,
                                   hipStream_t s)
{
    hipLaunchKernelGGL((k_cells_ref_slow_batched<8, 2, 512>), dim3(1024), dim3(64), 0, s, ins, rt,
                       want_density);
}

void launch_cells_ref(const BlockIn &in, const GridP &gp, const CellOut &o, int want_density,
                      hipStream_t s)
{
    if (in.ncells <= 0) return;
    hipLaunchKernelGGL((k_cells_ref<1, 1, 64>), dim3(xcd_grid(in.ncells)), dim3(64), 0, s, in,
                       gp, o, want_density);
}

void launch_cells_ref_slow(const BlockIn &in, const GridP &gp, const CellOut &o,
                           int want_density, hipStream_t s)
{
    if (in.ncells <= 0) return;
    hipLaunchKernelGGL((k_cells_ref_slow<8, 2, 512>), dim3(1024), dim3(64), 0, s, in, gp,
                       o, want_density);
}

// ---------------------------------------------------------------------------
// DENSE_CIC estimator: IterateCellsCic, src/dense.cpp:486-562
// ---------------------------------------------------------------------------
__global__ void k_cic(BlockIn in, GridP gp, CellOut o, const Router *rt, int self)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= in.ncells) return;
    const float4 p = in.pts[v];
    int mn[3];
    float m8[8];
    cic_weights(p.x, p.y, p.z, gp.mass, gp, mn, m8);
    for (int k = 0; k < 8; ++k)
        deposit_point(mn[0] + (k & 1), mn[1] + ((k >> 1) & 1), mn[2] + (k >> 2), m8[k],
                      cell_key(v, gp.project ? (k >> 2) : 0), gp, o, rt, self);
    o.status[v] = ST_CIC;
    o.npts[v] = 8;
}

void launch_cic(const BlockIn &in, const GridP &gp, const CellOut &o, const Router *rt_dev,
                int self, hipStream_t s)
{
    if (in.ncells <= 0) return;
    hipLaunchKernelGGL(k_cic, dim3((in.ncells + 255) / 256), dim3(256), 0, s, in, gp, o, rt_dev,
                       self);
}

}  // namespace t2
