// kernels.hpp -- device data layout and kernel declarations.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace t2 {

// Grid and block parameters of one dense run on one block (host-computed
// with the reference's GridStepParams / BlockGridParams arithmetic).
struct GridP {
    float gmin[3];
    float step[3];
    float dmin[3];
    float dmax[3];
    float bmin[3];       // block bounds (the in-block deposit test)
    float bmax[3];
    int bmin_idx[3];
    int bnum_idx[3];
    float eps;
    float mass;
    float div;
    int project;         // 2-D x-y projection
};

// counters in device memory (ctrs[])
enum {
    C_OVF = 0,        // entries appended to the shared-point list
    C_OFFBLOCK = 1,   // points outside the block bounds (dropped)
    C_SLOW = 2,       // cells queued for the large-capacity pass
    C_ERR = 3,        // bit 0: deferred-cell list full; bit 2: cell beyond all capacities
    C_OFFGRID = 4,    // in-bounds points whose index left the block grid
    C_FALLBACK = 5,   // FP64 re-checks of the exact assignment
    C_NCTR = 8
};

// Per-block device outputs of the density kernels.
struct CellOut {
    float *vol;        // [ncells] volume() or nullptr
    int *status;       // [ncells] TESS2_CELL_*
    int *npts;         // [ncells] points the cell deposits
    float *density;    // [grid]
    unsigned *count;   // [grid] contributions per point
    uint2 *first;      // [grid] first contribution (cell, mass bits)
    int *head;         // [grid] newest shared-list entry of the point
    int4 *ovf;         // shared-point list: {lidx, cell, mass bits, next}
    unsigned ovf_cap;
    unsigned *ctrs;    // [C_NCTR]
    int *slow;         // cells for the large-capacity pass
    unsigned slow_cap;
};

// uniform bucket hash of the particles: bucket (i,j,k) holds the particles
// with floor((x - lo) * inv) == (i,j,k), clamped to the grid of dims buckets
struct HashP {
    float lo[3];
    float inv;      // 1 / cell
    float cell;
    int dims[3];
};

struct BlockIn {
    const float4 *pts;   // xyz + pad
    const int4 *tets;    // 2 int4 per tet: verts, neighbours
    const int *v2t;
    const float4 *cc;    // circumcenters (xyz + pad)
    const int *order;    // cells in processing (Morton) order, or nullptr
    int ncells;          // original particles
    int ntets;
};

// Workgroup -> work item with XCD affinity: the dispatcher hands consecutive
// workgroups to the 8 XCDs in turn, so item k = (b % 8) * chunk + b / 8 gives
// each XCD (and its L2) one contiguous range of items.  Returns -1 for the
// padding workgroups of the last range.
__device__ __forceinline__ int xcd_item(int b, int n)
{
    const int chunk = (n + 7) >> 3;
    const int k = (b & 7) * chunk + (b >> 3);
    return k < n ? k : -1;
}
__host__ __forceinline__ int xcd_grid(int n) { return ((n + 7) >> 3) << 3; }

// circumcenters; cc.w = error bound of the float solve (inf when undefined)
void launch_circumcenters(const float4 *pts, const int4 *tets, int ntets, float4 *cc,
                          hipStream_t s);
void launch_hash_build(const float4 *pts, int np, const HashP &h, unsigned *start,
                       unsigned *fill, unsigned *sums, int *bk, float4 *sorted, hipStream_t s);
void launch_owner(const GridP &gp, const HashP &h, const unsigned *start, const float4 *hp,
                  int *owner, int prec, unsigned *ctrs, const int *status, int *npts,
                  int ncells, hipStream_t s);
void launch_exact_status(const BlockIn &in, const GridP &gp, int *status, int prec,
                         unsigned *ctrs, int *slow, unsigned slow_cap, hipStream_t s);
void launch_exact_tail(const BlockIn &in, const GridP &gp, const int *owner,
                       const int *status_ex, const CellOut &o, int prec, hipStream_t s);
void launch_fill_u32(unsigned *p, size_t n, unsigned v, hipStream_t s);
void launch_cells_ref(const BlockIn &in, const GridP &gp, const CellOut &o,
                      int want_density, hipStream_t s);
void launch_cells_ref_slow(const BlockIn &in, const GridP &gp, const CellOut &o,
                           int want_density, hipStream_t s);
void launch_cic(const BlockIn &in, const GridP &gp, const CellOut &o, hipStream_t s);
void launch_finalize(const CellOut &o, const GridP &gp, hipStream_t s);
void launch_cell_stats(const int *status, const int *npts, int ncells,
                       unsigned long long *out6, hipStream_t s);
void launch_grid_stats(const float *density, size_t n, double *sum, unsigned *maxbits,
                       hipStream_t s);

}  // namespace t2
