// ghosts.hip -- ghost selection of the tessellation rounds: the test of
// incomplete_cells() (src/tess.cpp:492-628), one lane per tet.  For every
// tet with an original (local) vertex: its circumsphere (float, as
// circumcenter() / distance(), src/tet.cpp:37-66,146-153) against each
// neighbour box, skipped when the sphere lies deep inside the block, and its
// convex-hull facets against each neighbour box (side_of_plane,
// src/tet.cpp:70-119).  The original vertices concerned get the neighbour's
// bit in a 64-bit mask (atomicOr); the host sends them as ghosts.
#include "kernels.hpp"
#include "geom.hpp"

namespace t2 {

namespace {

// dot(u, p, x) = u . (p - x) accumulated from 0, src/tet.cpp:21-29
__device__ __forceinline__ float dot3(const float *u, const float *p, const float *x)
{
    float r = 0.0f;
    r += u[0] * (p[0] - x[0]);
    r += u[1] * (p[1] - x[1]);
    r += u[2] * (p[2] - x[2]);
    return r;
}

// side_of_plane(), src/tet.cpp:70-119; q[4] holds the tet's vertices
__device__ bool side_of_plane(const float *bmin, const float *bmax, const float (*q)[3], int j)
{
    int idx[4] = {0, 1, 2, 3};
    idx[j] = 3;
    idx[3] = j;
    const float *x = q[idx[0]], *y = q[idx[1]], *z = q[idx[2]], *p = q[idx[3]];
    float u[3], v[3], n[3];
    for (int i = 0; i < 3; ++i) {
        u[i] = y[i] - x[i];
        v[i] = z[i] - x[i];
    }
    n[0] = u[1] * v[2] - u[2] * v[1];
    n[1] = u[2] * v[0] - u[0] * v[2];
    n[2] = u[0] * v[1] - u[1] * v[0];
    float res = dot3(n, p, x);
    const int sign = res > 0.0f ? 1 : (res < 0.0f ? -1 : 0);
    if (sign == 0) return true;   // flat hull tet: count the box as reached
    res = 0.0f;
    for (int i = 0; i < 3; ++i) {
        if ((sign > 0) != (n[i] < 0.0f)) res += n[i] * (bmin[i] - x[i]);
        else res += n[i] * (bmax[i] - x[i]);
    }
    return (sign > 0) != (res > 0.0f);
}

// distance from a point to a box (0 inside), in the float steps of DIY's
// distance(Bounds, Point) that src/tess.cpp:568 calls
__device__ __forceinline__ float box_distance(const float *bmin, const float *bmax,
                                              const float *p)
{
    float res = 0.0f;
    for (int i = 0; i < 3; ++i) {
        float diff = 0.0f, d;
        d = bmin[i] - p[i];
        if (d > diff) diff = d;
        d = p[i] - bmax[i];
        if (d > diff) diff = d;
        res += diff * diff;
    }
    return sqrtf(res);
}

}  // namespace

// boxes: 6 floats per neighbour (min xyz, max xyz), neighbours [first, nnb)
__global__ __launch_bounds__(256) void k_ghost_sends(const float4 *__restrict__ pts,
                                                     const int4 *__restrict__ tets, int ntets,
                                                     int norig, float3 bmin, float3 bmax,
                                                     const float *__restrict__ boxes, int nnb,
                                                     int first,
                                                     unsigned long long *__restrict__ sends)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntets) return;
    const int4 vv = tets[2 * t];
    const int vid[4] = {vv.x, vv.y, vv.z, vv.w};
    if (vv.x >= norig && vv.y >= norig && vv.z >= norig && vv.w >= norig) return;
    const int4 nn = tets[2 * t + 1];
    const int nid[4] = {nn.x, nn.y, nn.z, nn.w};
    float q[4][3];
    for (int k = 0; k < 4; ++k) {
        const float4 a = pts[vid[k]];
        q[k][0] = a.x;
        q[k][1] = a.y;
        q[k][2] = a.z;
    }
    float c[3];
    circumcenter(q[0][0], q[0][1], q[0][2], q[1][0], q[1][1], q[1][2], q[2][0], q[2][1],
                 q[2][2], q[3][0], q[3][1], q[3][2], c[0], c[1], c[2]);
    // distance(center, first vertex): sum of squares from 0, then the root
    float r2 = 0.0f;
    for (int i = 0; i < 3; ++i) r2 += (c[i] - q[0][i]) * (c[i] - q[0][i]);
    const float rad = sqrtf(r2);
    unsigned long long vbits[4] = {0ull, 0ull, 0ull, 0ull};
    // convex hull facets that carry a local vertex (:516-547)
    for (int j = 0; j < 4; ++j) {
        if (nid[j] != -1) continue;
        bool local = false;
        for (int k = 0; k < 4; ++k)
            if (k != j && vid[k] < norig) local = true;
        if (!local) continue;
        unsigned long long m = 0ull;
        for (int i = first; i < nnb; ++i)
            if (side_of_plane(&boxes[6 * i], &boxes[6 * i + 3], q, j)) m |= 1ull << i;
        for (int k = 0; k < 4; ++k)
            if (k != j) vbits[k] |= m;
    }
    // a circumsphere deep inside the block cannot reach a neighbour (:550-557)
    const float bl[3] = {bmin.x, bmin.y, bmin.z}, bh[3] = {bmax.x, bmax.y, bmax.z};
    bool deep = true;
    for (int j = 0; j < 3; ++j) {
        if (c[j] - bl[j] <= rad || bh[j] - c[j] <= rad) {
            deep = false;
            break;
        }
    }
    if (!deep) {
        unsigned long long m = 0ull;
        for (int i = first; i < nnb; ++i)
            if (box_distance(&boxes[6 * i], &boxes[6 * i + 3], c) <= rad) m |= 1ull << i;
        for (int k = 0; k < 4; ++k) vbits[k] |= m;
    }
    for (int k = 0; k < 4; ++k)
        if (vid[k] < norig && vbits[k]) atomicOr(&sends[vid[k]], vbits[k]);
}

void launch_ghost_sends(const float4 *pts, const int4 *tets, int ntets, int norig,
                        const float *bmin, const float *bmax, const float *boxes, int nnb,
                        int first, unsigned long long *sends, hipStream_t s)
{
    if (ntets <= 0) return;
    hipLaunchKernelGGL(k_ghost_sends, dim3((ntets + 255) / 256), dim3(256), 0, s, pts, tets,
                       ntets, norig, make_float3(bmin[0], bmin[1], bmin[2]),
                       make_float3(bmax[0], bmax[1], bmax[2]), boxes, nnb, first, sends);
}

}  // namespace t2
