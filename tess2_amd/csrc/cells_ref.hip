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

namespace t2 {

enum { ST_SCAN = 0, ST_CIC = 1, ST_INCOMPLETE = 2, ST_OUTSIDE = 3, ST_DEGENERATE = 4 };

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

__device__ __forceinline__ bool wide_bit(unsigned long long w, int b) { return (w >> b) & 1ull; }

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

template <int NS, int NF, int RMAX, bool WIDE = false>
__device__ void cell_ref_body(const int v, const BlockIn &in, const GridP &gp,
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
                if (slot < SMAX) s_tid[slot] = cand;
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
        // gather the records of the new slots [f1, S)
        for (int b = f1; b < S; b += 64) {
            const int s = b + lane;
            if (s < S) {
                const int id = s_tid[s];
                s_v[s] = in.tets[2 * id];
                s_u[s] = in.tets[2 * id + 1];
                s_cc[s] = in.cc[id];
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
        // probes.  Larger boxes probe face-parallel as they go.
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
        unsigned long long mk = 0ull;
        unsigned long long w1 = 0ull, w2 = 0ull, w3 = 0ull, w4 = 0ull, w5 = 0ull, w6 = 0ull,
                           w7 = 0ull;
        float4 *s_fa = s_cc;   // star records and circumcenters are done
        float2 *s_fb = reinterpret_cast<float2 *>(s_v);
        if (use_mask || use_wide) {
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
                const unsigned long long m = __ballot(p < nbox && !(hi >= 0.0f && lo < 0.0f));
                if (c == 0) mk = m;
                else if (c == 1) w1 = m;
                else if (c == 2) w2 = m;
                else if (c == 3) w3 = m;
                else if (c == 4) w4 = m;
                else if (c == 5) w5 = m;
                else if (c == 6) w6 = m;
                else w7 = m;
            }
        }
        if (use_mask) {
            const int x = lane & ((1 << lb0) - 1), y = (lane >> lb0) & ((1 << lb1) - 1);
            const int z = lane >> lb01;
            const float qx = p0 + (float)x * gp.step[0];
            const float qy = p1 + (float)y * gp.step[1];
            const float qz = p2 + (float)z * gp.step[2];
            // largest and smallest significant distance: a face with d >= 0
            // exists iff hi >= 0, one with d < 0 iff lo < 0
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
            const bool in_cell = !(hi >= 0.0f && lo < 0.0f);
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
    ((k) == 0 ? mk : (k) == 1 ? w1 : (k) == 2 ? w2 : (k) == 3 ? w3 : (k) == 4 ? w4 : (k) == 5 ? w5 \
     : (k) == 6 ? w6 : w7)
#define PROBE(xi, yi, zi)                                                                     \
    (use_mask ? (bool)((mk >> ((xi) | ((yi) << lb0) | ((zi) << lb01))) & 1ull)                 \
     : use_wide ? wide_bit(WORD(((xi) + n0 * ((yi) + n1 * (zi))) >> 6),                        \
                           ((xi) + n0 * ((yi) + n1 * (zi))) & 63)                               \
              : inside(p0 + (float)(xi) * gp.step[0], p1 + (float)(yi) * gp.step[1],             \
                       p2 + (float)(zi) * gp.step[2]))

        // ---- 5. scanline search (src/dense.cpp:1486-1676) --------------------
        // Pass 0 counts the points and keeps the runs (one per z-y line) in
        // LDS.  A cell with more than RMAX runs scans again in pass 1 -- the
        // state machine takes the same path -- and deposits each run as it is
        // found, now that mass/N is known.
        int4 *runs = s_u;   // star neighbour records are no longer needed
        int nruns = 0, tot = 0;
        float mpp = 0.0f;
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
                        } else {
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
            if (lane == 0) {   // deposited by pass 1
                o.status[v] = ST_SCAN;
                o.npts[v] = tot;
            }
        } else if (tot > 0) {
            mpp = gp.mass / (float)tot;   // src/dense.cpp:1692
            for (int base = 0; base < tot; base += 64) {
                const int p = base + lane;
                int r = 0;
                for (int k = 1; k < nruns; ++k)
                    if (p >= runs[k].w) r = k;
                if (p < tot) {
                    const int4 rr = runs[r];
                    deposit_point(cmi0 + rr.z + (p - rr.w), cmi1 + rr.y, cmi2 + rr.x, mpp,
                                  cell_key(v, 0), gp, o, rt, self);
                }
            }
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
    cell_ref_body<NS, NF, RMAX>(v, in, gp, o, want_density, true, s_v, s_ns, s_cc, s_u, s_h);
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
        cell_ref_body<NS, NF, RMAX>(v, in, gp, o, want_density, false, s_v, s_ns, s_cc, s_u, s_h);
        __syncthreads();
    }
}

// All blocks of a run in one persistent launch: every wave takes chunks of
// cells (one block each, consecutive in the block's Morton order) from the
// queue of its XCD's range of the chunk list, then from the other ranges,
// until all are empty.  Chunk sizes by block size class (tess2_autotune).  At least 6
// waves per SIMD: the per-block descriptors read from memory otherwise push
// the register count down to 4 waves (30 vs 34.5 ms on 64 Plummer blocks).
template <int NS, int NF, int RMAX>
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
        const GridP gp = gps[ch.block];
        const CellOut o = outs[ch.block];
        for (int j = 0; j < ch.count; ++j) {
            const int k = ch.start + j;
            const int v = in.order ? in.order[k] : k;
            cell_ref_body<NS, NF, RMAX, true>(v, in, gp, o, want_density, true, s_v, s_ns, s_cc, s_u,
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
    const int nb = rt->nblocks;
    for (int b = 0; b < nb; ++b) {
        const CellOut &o = rt->outs[b];
        unsigned n = o.ctrs[C_SLOW];
        if (n > o.slow_cap) n = o.slow_cap;
        for (unsigned k = blockIdx.x; k < n; k += gridDim.x) {
            cell_ref_body<NS, NF, RMAX, true>(o.slow[k], ins[b], rt->gps[b], o, want_density,
                                              false, s_v, s_ns, s_cc, s_u, s_h, rt, b);
            __syncthreads();
        }
    }
}

void launch_cells_ref_batched(const BlockIn *ins, const GridP *gps, const CellOut *outs,
                              const Router *rt, const Chunk *chunks, const XcdQueues &xq,
                              unsigned *queue, int waves, int want_density, hipStream_t s)
{
    launch_fill_u32(queue, 8, 0u, s);
    if (xq.start[8] > 0)
        hipLaunchKernelGGL((k_cells_ref_batched<1, 1, 64>), dim3(waves), dim3(64), 0, s, ins, gps,
                           outs, rt, chunks, xq, queue, want_density);
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
