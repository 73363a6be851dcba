// deposit.hpp -- the grid deposit shared by the estimators: in-block test,
// local index, and the ordered-accumulation protocol (first contribution
// written directly, later ones listed for the ordered fold).
#pragma once

#include "geom.hpp"
#include "kernels.hpp"

namespace t2 {

// ---------------------------------------------------------------------------
// ordered-accumulation protocol on one block grid point: the first
// contribution writes the density; later ones join the point's list.  The
// key orders the fold: a block's own cells by cell id, then points received
// from other blocks by (source rank, cell) (bit 31 set).  Its low bit is the
// z offset of a CIC corner in the 2-D projection, where corners z and z+1 of
// one cell land on the same (x, y) point: the reference adds them in corner
// order (z first).  Scan points of one cell that share (x, y) all carry the
// same mass, so their order cannot change the sum.
//   own:      cell << 1 | dz                         (cells < 2^30)
//   received: 1 << 31 | rank << 24 | cell << 1 | dz  (cells < 2^23, rank < 127)
// ---------------------------------------------------------------------------
__device__ __forceinline__ unsigned cell_key(int cell, int dz)
{
    return ((unsigned)cell << 1) | (unsigned)dz;
}

__device__ __forceinline__ void push_point(long long lidx, float m, unsigned key, float div,
                                           const CellOut &o)
{
    const unsigned c = atomicAdd(&o.count[lidx], 1u);
    if (c == 0u) {
        o.first[lidx] = make_uint2(key, __float_as_uint(m));
        o.density[lidx] = (float)((double)m / (double)div);
    } else if (c == 1u) {
        // most shared points have two contributions: the second sits next to
        // the point too, with no list slot (no second dependent atomic)
        o.second[lidx] = make_uint2(key, __float_as_uint(m));
    } else {
        const unsigned e = atomicAdd(&o.ctrs[C_OVF], 1u);
        // the entry keeps its ordinal c among the point's contributions: the
        // entry with c == 2 represents the point in the fold, and the fold's
        // scatter puts each entry at slot c - 2 of the point's segment
        if (e < o.ovf_cap) o.ovf[e] = make_int4((int)lidx, (int)key, __float_as_int(m), (int)c);
    }
}

// local index of a global grid index in a block's grid, -1 outside it.  The
// 2-D projection drops z from the index (index(), src/dense.cpp:1076-1089),
// so a point of the block's own cells only needs its x and y in range (its z
// is within the block's bounds already); a point routed to another block must
// lie in that block's 3-D grid range (holds3d), as the reference sends it to
// the neighbour whose bounds contain it in 3-D.
__device__ __forceinline__ long long block_lidx(int gx, int gy, int gz, const GridP &gp,
                                                bool holds3d = false)
{
    const int lx = gx - gp.bmin_idx[0], ly = gy - gp.bmin_idx[1], lz = gz - gp.bmin_idx[2];
    if (lx < 0 || ly < 0 || lx >= gp.bnum_idx[0] || ly >= gp.bnum_idx[1]) return -1;
    if (gp.project && !holds3d) return (long long)ly * gp.bnum_idx[0] + lx;
    if (lz < 0 || lz >= gp.bnum_idx[2]) return -1;
    if (gp.project) return (long long)ly * gp.bnum_idx[0] + lx;
    return ((long long)lz * gp.bnum_idx[1] + ly) * gp.bnum_idx[0] + lx;
}

// ---------------------------------------------------------------------------
// deposit one grid point (src/dense.cpp:276-305).  In the block (inside its
// bounds and its grid): the block's own grid.  Otherwise the point goes to
// the block whose grid holds its global index, as received mass
// (src/dense.cpp:298-305 -> recvd_pts :166-202), or is dropped and counted.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void deposit_point(int gx, int gy, int gz, float m, unsigned ckey,
                                              const GridP &gp, const CellOut &o,
                                              const Router *rt = nullptr, int self = 0)
{
    const float px = idx2phys1(gx, gp.step[0], gp.gmin[0]);
    const float py = idx2phys1(gy, gp.step[1], gp.gmin[1]);
    const float pz = idx2phys1(gz, gp.step[2], gp.gmin[2]);
    const long long lidx = block_lidx(gx, gy, gz, gp);
    if (lidx >= 0 && px >= gp.bmin[0] && px <= gp.bmax[0] && py >= gp.bmin[1] &&
        py <= gp.bmax[1] && pz >= gp.bmin[2] && pz <= gp.bmax[2]) {
        push_point(lidx, m, ckey, gp.div, o);
        return;
    }
    if (rt) {
        // the first block other than this one whose 3-D grid range holds
        // the point: the first of the two candidates of its interval box
        // that is not this block (two table loads, no walk over the blocks)
        const int ix = gx - rt->rbase[0], iy = gy - rt->rbase[1], iz = gz - rt->rbase[2];
        if (ix >= 0 && iy >= 0 && iz >= 0 && ix < rt->rlen[0] && iy < rt->rlen[1] &&
            iz < rt->rlen[2]) {
            const int a = rt->rmap[rt->roff[0] + ix], b = rt->rmap[rt->roff[1] + iy],
                      c = rt->rmap[rt->roff[2] + iz];
            const int2 cd = rt->rcand[((long long)a * rt->rn[1] + b) * rt->rn[2] + c];
            const int j = cd.x != self ? cd.x : cd.y;
            if (j >= 0) {
                const long long lj = block_lidx(gx, gy, gz, rt->gps[j], true);
                const unsigned key = 0x80000000u | ((unsigned)self << 24) | ckey;
                push_point(lj, m, key, rt->gps[j].div, rt->outs[j]);
                atomicAdd(&o.ctrs[C_ROUTED], 1u);
                return;
            }
        }
    }
    atomicAdd(&o.ctrs[C_OFFBLOCK], 1u);
}

// DistributeScalarCIC(): src/dense.cpp:1787-1872, corner k in z,y,x order.
__device__ void cic_weights(float sx, float sy, float sz, float scalar, const GridP &gp,
                            int mn[3], float m8[8])
{
    mn[0] = phys2idx1(sx, gp.step[0], gp.gmin[0]);
    mn[1] = phys2idx1(sy, gp.step[1], gp.gmin[1]);
    mn[2] = phys2idx1(sz, gp.step[2], gp.gmin[2]);
    float tot = 0.0f, v0 = 0.0f;
    const float e2 = 2.0f * gp.eps;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int ix = mn[0] + (k & 1), iy = mn[1] + ((k >> 1) & 1), iz = mn[2] + (k >> 2);
        const float gx = idx2phys1(ix, gp.step[0], gp.gmin[0]);
        const float gy = idx2phys1(iy, gp.step[1], gp.gmin[1]);
        const float gz = idx2phys1(iz, gp.step[2], gp.gmin[2]);
        float px = sx, py = sy, pz = sz;
        if (fabsf(px - gx) < gp.eps) px += e2;
        if (fabsf(py - gy) < gp.eps) py += e2;
        if (fabsf(pz - gz) < gp.eps) pz += e2;
        const float vol = fabsf((gx - px) * (gy - py) * (gz - pz));
        if (v0 == 0.0f) v0 = vol;
        const float w = v0 / vol;
        m8[k] = w;
        tot += w;
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const float w = m8[k] / tot;
        m8[k] = w * scalar;
    }
}

}  // namespace t2
