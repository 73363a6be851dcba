// common.hip -- streaming kernels: circumcenters, fills, statistics.
#include "kernels.hpp"
#include "geom.hpp"

#include <algorithm>

namespace t2 {

// fill_circumcenters(): one lane per tet, src/volume.cpp:6-11.  The four
// vertex gathers are 16-byte loads of the padded particle array.
__global__ void k_circumcenters(const float4 *__restrict__ pts, const int4 *__restrict__ tets,
                                int ntets, float4 *__restrict__ cc)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntets) return;
    const int4 vv = tets[2 * t];
    const float4 a = pts[vv.x], b = pts[vv.y], c = pts[vv.z], d = pts[vv.w];
    float x, y, z;
    circumcenter(a.x, a.y, a.z, b.x, b.y, b.z, c.x, c.y, c.z, d.x, d.y, d.z, x, y, z);
    const float e = circumcenter_err(a.x, a.y, a.z, b.x, b.y, b.z, c.x, c.y, c.z, d.x, d.y, d.z,
                                     x, y, z);
    cc[t] = make_float4(x, y, z, e);
}

void launch_circumcenters(const float4 *pts, const int4 *tets, int ntets, float4 *cc,
                          hipStream_t s)
{
    if (ntets <= 0) return;
    hipLaunchKernelGGL(k_circumcenters, dim3((ntets + 255) / 256), dim3(256), 0, s, pts, tets,
                       ntets, cc);
}

// circumcenters of every staged block, block = blockIdx.y
__global__ void k_circumcenters_blocks(const BlockIn *__restrict__ ins)
{
    const BlockIn &in = ins[blockIdx.y];
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < in.ntets;
         t += gridDim.x * blockDim.x) {
        const int4 vv = in.tets[2 * t];
        const float4 a = in.pts[vv.x], b = in.pts[vv.y], c = in.pts[vv.z], d = in.pts[vv.w];
        float x, y, z;
        circumcenter(a.x, a.y, a.z, b.x, b.y, b.z, c.x, c.y, c.z, d.x, d.y, d.z, x, y, z);
        const float e = circumcenter_err(a.x, a.y, a.z, b.x, b.y, b.z, c.x, c.y, c.z, d.x, d.y,
                                         d.z, x, y, z);
        const_cast<float4 *>(in.cc)[t] = make_float4(x, y, z, e);   // read-only downstream
    }
}

// zero the density and count grids of every block, block = blockIdx.y
__global__ void k_clear_grids(const CellOut *__restrict__ outs)
{
    const CellOut &o = outs[blockIdx.y];
    const unsigned long long n = o.ngrid, n4 = n / 4;
    const unsigned long long stride = (unsigned long long)gridDim.x * blockDim.x;
    const unsigned long long i0 = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
    const uint4 z = make_uint4(0u, 0u, 0u, 0u);
    for (unsigned long long i = i0; i < n4; i += stride) {
        reinterpret_cast<uint4 *>(o.density)[i] = z;
        reinterpret_cast<uint4 *>(o.count)[i] = z;
    }
    if (i0 < n - n4 * 4) {
        o.density[n4 * 4 + i0] = 0.0f;
        o.count[n4 * 4 + i0] = 0u;
    }
}

void launch_block_prep(const BlockIn *ins, const CellOut *outs, int nb, int max_tets,
                       unsigned long long max_grid, bool circumcenters, hipStream_t s)
{
    if (nb <= 0) return;
    unsigned bg = (unsigned)std::min<unsigned long long>(2048, (max_grid / 4 + 255) / 256);
    if (bg == 0) bg = 1;
    hipLaunchKernelGGL(k_clear_grids, dim3(bg, nb), dim3(256), 0, s, outs);
    if (circumcenters && max_tets > 0) {
        const unsigned bt = (unsigned)std::min(16384, (max_tets + 255) / 256);
        hipLaunchKernelGGL(k_circumcenters_blocks, dim3(bt, nb), dim3(256), 0, s, ins);
    }
}

__global__ void k_fill_u32(uint4 *__restrict__ p4, size_t n4, unsigned *__restrict__ tail,
                           size_t ntail, unsigned v)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    const uint4 q = make_uint4(v, v, v, v);
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride)
        p4[i] = q;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < ntail) tail[i] = v;
}

// fill a 16-byte aligned buffer (a kernel rather than a memset so that the
// sequence stays valid inside a captured graph)
void launch_fill_u32(unsigned *p, size_t n, unsigned v, hipStream_t s)
{
    if (n == 0) return;
    const size_t n4 = n / 4, ntail = n - n4 * 4;
    size_t blocks = (n4 + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    if (blocks == 0) blocks = 1;
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)blocks), dim3(256), 0, s, (uint4 *)p, n4,
                       p + n4 * 4, ntail, v);
}

// particles given as xyz triples -> the padded float4 layout of the kernels
__global__ void k_pack_points(const float *__restrict__ xyz, int np, float4 *__restrict__ pts)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    pts[p] = make_float4(xyz[3 * p], xyz[3 * p + 1], xyz[3 * p + 2], 0.0f);
}

void launch_pack_points(const float *xyz, int np, float4 *pts, hipStream_t s)
{
    if (np <= 0) return;
    hipLaunchKernelGGL(k_pack_points, dim3((np + 255) / 256), dim3(256), 0, s, xyz, np, pts);
}

// index ranges of a block held in device memory: every tet vertex in
// [0, np), every neighbour in [-1, nt), every vert_to_tet entry in [-1, nt);
// a violation sets bit `bit` of *err
__global__ void k_check_block(const int4 *__restrict__ tets, int nt, const int *__restrict__ v2t,
                              int np, unsigned *__restrict__ err, unsigned bit)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    bool bad = false;
    if (i < nt) {
        const int4 v = tets[2 * i], n = tets[2 * i + 1];
        bad = (unsigned)v.x >= (unsigned)np || (unsigned)v.y >= (unsigned)np ||
              (unsigned)v.z >= (unsigned)np || (unsigned)v.w >= (unsigned)np ||
              n.x < -1 || n.x >= nt || n.y < -1 || n.y >= nt || n.z < -1 || n.z >= nt ||
              n.w < -1 || n.w >= nt;
    }
    if (i < np) bad = bad || v2t[i] < -1 || v2t[i] >= nt;
    if (bad) atomicOr(err, bit);
}

void launch_check_block(const int4 *tets, int nt, const int *v2t, int np, unsigned *err,
                        unsigned bit, hipStream_t s)
{
    const int n = nt > np ? nt : np;
    if (n <= 0) return;
    hipLaunchKernelGGL(k_check_block, dim3((n + 255) / 256), dim3(256), 0, s, tets, nt, v2t, np,
                       err, bit);
}

// per block (blockIdx.y): status histogram (5 codes) and total points of
// deposited cells
__global__ __launch_bounds__(256) void k_cell_stats(const CellOut *__restrict__ outs)
{
    // cells per status counted by ballots (wave-uniform counters), points of
    // deposited cells summed per lane; one atomic per counter per workgroup
    __shared__ unsigned long long s_c[4][6];
    const CellOut &o = outs[blockIdx.y];
    unsigned long long c[5] = {0, 0, 0, 0, 0}, np = 0;
    for (int base = blockIdx.x * blockDim.x; base < o.ncells; base += gridDim.x * blockDim.x) {
        const int i = base + threadIdx.x;
        int s = -1;
        if (i < o.ncells) {
            s = o.status[i];
            if (s == 0 || s == 1) np += (unsigned long long)o.npts[i];
        }
#pragma unroll
        for (int k = 0; k < 5; ++k) c[k] += (unsigned long long)__popcll(__ballot(s == k));
    }
    for (int off = 32; off >= 1; off >>= 1) np += __shfl_xor(np, off);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int k = 0; k < 5; ++k) s_c[w][k] = c[k];
        s_c[w][5] = np;
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        const int k = threadIdx.x;
        const unsigned long long x = s_c[0][k] + s_c[1][k] + s_c[2][k] + s_c[3][k];
        if (x) atomicAdd(&o.st->cstat[k], x);
    }
}

// per block (blockIdx.y): sum (double) and max of the density grid.  The
// sum is reduced in a fixed order (lanes, waves, then workgroups in
// k_grid_final) so that it is the same on every run.
constexpr int GRID_PARTS = 1024;

__global__ __launch_bounds__(256) void k_grid_stats(const CellOut *__restrict__ outs,
                                                    double *__restrict__ part)
{
    __shared__ double s_acc[4];
    const CellOut &o = outs[blockIdx.y];
    // 16-byte loads, four independent partial sums per lane (the grid
    // arrays are 256-byte aligned allocations); the last n mod 4 points
    // go to the first lanes of workgroup 0
    const unsigned long long n = o.ngrid, n4 = n >> 2;
    const float4 *d4 = reinterpret_cast<const float4 *>(o.density);
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    float mx = 0.0f;
#pragma unroll 4
    for (unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; i < n4;
         i += (unsigned long long)gridDim.x * blockDim.x) {
        const float4 x = d4[i];
        a0 += (double)x.x;
        a1 += (double)x.y;
        a2 += (double)x.z;
        a3 += (double)x.w;
        mx = fmaxf(mx, fmaxf(fmaxf(x.x, x.y), fmaxf(x.z, x.w)));
    }
    if (blockIdx.x == 0 && threadIdx.x < (unsigned)(n & 3ull)) {
        const float x = o.density[(n4 << 2) + threadIdx.x];
        a0 += (double)x;
        mx = fmaxf(mx, x);
    }
    double acc = (a0 + a1) + (a2 + a3);
    for (int off = 32; off >= 1; off >>= 1) {
        acc += __shfl_xor(acc, off);
        mx = fmaxf(mx, __shfl_xor(mx, off));
    }
    if ((threadIdx.x & 63) == 0) {
        s_acc[threadIdx.x >> 6] = acc;
        if (mx != 0.0f) atomicMax(&o.st->gmax, __float_as_uint(mx));   // order-free
    }
    __syncthreads();
    if (threadIdx.x == 0)
        part[blockIdx.y * GRID_PARTS + blockIdx.x] = ((s_acc[0] + s_acc[1]) + s_acc[2]) + s_acc[3];
}

__global__ __launch_bounds__(256) void k_grid_final(const CellOut *__restrict__ outs,
                                                    const double *__restrict__ part, int nparts)
{
    __shared__ double s[GRID_PARTS];
    for (int k = threadIdx.x; k < nparts; k += blockDim.x) s[k] = part[blockIdx.x * GRID_PARTS + k];
    __syncthreads();
    if (threadIdx.x != 0) return;
    double acc = 0.0;
    for (int k = 0; k < nparts; ++k) acc += s[k];
    outs[blockIdx.x].st->gsum = acc;
}

void launch_stats(const CellOut *outs, double *part, int nb, int max_cells,
                  unsigned long long max_grid, hipStream_t s)
{
    if (nb <= 0) return;
    // about 16 items per lane, at most GRID_PARTS workgroups per block (a
    // single block fills the chip; many blocks share it)
    const long long cap = std::max<long long>(16, GRID_PARTS / std::max(1, nb / 4));
    unsigned bc = (unsigned)std::min<long long>(cap, ((long long)max_cells + 4095) / 4096);
    unsigned bg = (unsigned)std::min<unsigned long long>(cap, (max_grid + 4095) / 4096);
    if (bc == 0) bc = 1;
    if (bg == 0) bg = 1;
    hipLaunchKernelGGL(k_cell_stats, dim3(bc, nb), dim3(256), 0, s, outs);
    hipLaunchKernelGGL(k_grid_stats, dim3(bg, nb), dim3(256), 0, s, outs, part);
    hipLaunchKernelGGL(k_grid_final, dim3(nb), dim3(256), 0, s, outs, (const double *)part,
                       (int)bg);
}

}  // namespace t2
