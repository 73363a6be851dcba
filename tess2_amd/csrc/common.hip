// common.hip -- streaming kernels: circumcenters, fills, statistics.
#include "kernels.hpp"
#include "geom.hpp"

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

// status histogram (5 codes) and total points of deposited cells
__global__ void k_cell_stats(const int *__restrict__ status, const int *__restrict__ npts,
                             int n, unsigned long long *out6)
{
    unsigned long long c[6] = {0, 0, 0, 0, 0, 0};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int s = status[i];
        if (s >= 0 && s < 5) c[s]++;
        if (s == 0 || s == 1) c[5] += (unsigned long long)npts[i];
    }
    for (int k = 0; k < 6; ++k) {
        unsigned long long x = c[k];
        for (int off = 32; off >= 1; off >>= 1) x += __shfl_xor(x, off);
        if ((threadIdx.x & 63) == 0 && x) atomicAdd(&out6[k], x);
    }
}

void launch_cell_stats(const int *status, const int *npts, int ncells,
                       unsigned long long *out6, hipStream_t s)
{
    if (ncells <= 0) return;
    int blocks = (ncells + 255) / 256;
    if (blocks > 1024) blocks = 1024;
    hipLaunchKernelGGL(k_cell_stats, dim3(blocks), dim3(256), 0, s, status, npts, ncells,
                       out6);
}

// sum (double) and max of the density grid
__global__ void k_grid_stats(const float *__restrict__ d, size_t n, double *sum,
                             unsigned *maxbits)
{
    double acc = 0.0;
    float mx = 0.0f;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        const float x = d[i];
        acc += (double)x;
        mx = fmaxf(mx, x);
    }
    for (int off = 32; off >= 1; off >>= 1) {
        acc += __shfl_xor(acc, off);
        mx = fmaxf(mx, __shfl_xor(mx, off));
    }
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(sum, acc);
        atomicMax(maxbits, __float_as_uint(mx));   // non-negative floats order as uints
    }
}

void launch_grid_stats(const float *density, size_t n, double *sum, unsigned *maxbits,
                       hipStream_t s)
{
    if (n == 0) return;
    size_t blocks = (n + 255) / 256;
    if (blocks > 1024) blocks = 1024;
    hipLaunchKernelGGL(k_grid_stats, dim3((unsigned)blocks), dim3(256), 0, s, density, n, sum,
                       maxbits);
}

}  // namespace t2
