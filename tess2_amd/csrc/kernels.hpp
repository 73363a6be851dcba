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
    C_OFFBLOCK = 1,   // points no block grid holds (dropped)
    C_SLOW = 2,       // cells queued for the large-capacity pass
    C_ERR = 3,        // bit 0: deferred-cell list full; bit 2: cell beyond all capacities;
                      // bit 3: a shared point beyond the fold's capacities
    C_DEFER = 4,      // grid points queued for the FP64 search of the exact assignment
    C_FALLBACK = 5,   // FP64 re-checks of the exact assignment
    C_ROUTED = 6,     // points deposited into another block's grid
    C_SEG = 7,        // contributions gathered into per-point segments for the fold
    C_FOLD0 = 8,      // shared points per fold class (C_FOLD0 + class, 3 classes)
    C_PAIRS = 11,     // points with a second contribution
    C_NCTR = 12
};

// fold classes by contributions per shared point
enum { FOLD_SMALL_MAX = 16, FOLD_MID_MAX = 2048 };

// Counters and statistics of one block; the records of all blocks of a
// context sit in one device array (cleared by one fill, copied to pinned host
// memory by one transfer).
struct BlockStats {
    unsigned ctrs[C_NCTR];
    unsigned pad0[16 - C_NCTR];
    unsigned long long cstat[6];   // cells per status 0..4, points of deposited cells
    double gsum;                   // sum of the block's density grid
    unsigned gmax;                 // bits of its maximum (non-negative float)
    unsigned pad1;
};
static_assert(sizeof(BlockStats) % 16 == 0, "BlockStats is cleared with 16-byte stores");

// Per-block device outputs of the density kernels.
struct CellOut {
    float *vol;        // [ncells] volume() or nullptr
    int *status;       // [ncells] TESS2_CELL_*
    int *npts;         // [ncells] points the cell deposits
    float *density;    // [grid]
    unsigned *count;   // [grid] contributions per point
    uint2 *first;      // [grid] first contribution (cell, mass bits)
    uint2 *second;     // [grid] second contribution (the list holds the third on)
    int *head;         // [grid] the point's fold record, -(1 + 4 index + class), written
                       // by the fold's gather for points with list entries
    int4 *ovf;         // shared-point list, third contribution on: {lidx, key, mass bits,
                       // ordinal among the point's contributions (2 for the third)};
                       // (3 * ovf_cap + 1) / 2 entries, the large fold's scratch
    unsigned ovf_cap;
    unsigned *ctrs;    // [C_NCTR]
    int *slow;         // cells for the large-capacity pass
    unsigned slow_cap;
    uint2 *seg;        // [3 * ovf_cap] contributions {key, mass bits}, one segment per point
    int4 *fold[3];     // shared points per fold class: {lidx, segment start, count, cursor}
    unsigned fold_cap[3];
    BlockStats *st;    // this block's record (st->ctrs == ctrs)
    unsigned long long ngrid;
    int ncells;
};

// uniform bucket hash of the particles: bucket (i,j,k) holds the particles
// with floor((x - lo) * inv) == (i,j,k), clamped to the grid of dims buckets
struct HashP {
    float lo[3];
    float inv;      // 1 / cell
    float cell;
    int dims[3];
};

// Per-block grids and outputs of all blocks of a run, in staging order (the
// rank of a block = its position; blocks are staged in ascending gid), for
// deposits that leave their block.
struct Router {
    int nblocks;
    const GridP *gps;
    const CellOut *outs;
    // The ends of the blocks' 3-D grid ranges cut each axis into intervals:
    // rmap[roff[a] + g - rbase[a]] is the interval of global index g on
    // axis a (g in [rbase[a], rbase[a] + rlen[a])), rcand[(i0 * rn[1] + i1)
    // * rn[2] + i2] the first two blocks (staging order, -1 for none) whose
    // ranges hold the box of intervals (i0, i1, i2).
    const int *rmap;
    const int2 *rcand;
    int rbase[3], rlen[3], roff[3], rn[3];
};

// one chunk of cells of one block for the batched launch
struct Chunk {
    int block;
    int start;    // position in the block's cell order of the first cell
    int count;
    int stride;   // cells start, start + stride, ... (count of them)
};

// The chunk list split into 8 contiguous ranges, one per XCD (workgroup b
// runs on XCD b % 8): waves take chunks of their own range first, so that
// each XCD's L2 sees one compact part of the data, then help the others.
struct XcdQueues {
    int start[9];   // range x = chunks [start[x], start[x+1])
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

// find or insert key (>= 0) in an open-addressing table of HS entries in
// LDS (-1 = empty); returns the entry, or -1 when HS probes found no room
template <int HS>
__device__ __forceinline__ int hash_find(int *hk, int key)
{
    unsigned h = ((unsigned)key * 0x9E3779B1u) >> 16;
    for (int n = 0; n < HS; ++n) {
        h &= HS - 1;
        const int old = atomicCAS(&hk[h], -1, key);
        if (old == -1 || old == key) return (int)h;
        ++h;
    }
    return -1;
}

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
                       unsigned *fill, unsigned *sums, int *bk, float4 *sorted, const int *key,
                       hipStream_t s);
// npts_tab / stex_tab (by rank) when several blocks share the assignment,
// else nullptr; defq: queue of defq_cap grid points for the FP64 re-search
// of the mixed precision (0: searched inline)
void launch_owner(const GridP &gp, const HashP &h, const unsigned *start, const float4 *hp,
                  int *owner, int prec, unsigned *ctrs, const int *status, int *npts,
                  int ncells, int *const *npts_tab, const int *const *stex_tab, unsigned *defq,
                  unsigned defq_cap, hipStream_t s);
// flags: 3 * ncells bytes of scratch (incomplete, outside, defined)
void launch_exact_status(const BlockIn &in, const GridP &gp, int *status, int prec,
                         unsigned *ctrs, unsigned char *flags, hipStream_t s);
// parts: 1 = densities of the owned points, 2 = CIC of cells that own none;
// every block's part 1 must precede any block's part 2
void launch_exact_tail(const BlockIn &in, const GridP &gp, const int *owner,
                       const int *status_ex, const CellOut &o, int prec, int *const *npts_tab,
                       const int *const *stex_tab, const Router *rt, int self, int parts,
                       hipStream_t s);
void launch_fill_u32(unsigned *p, size_t n, unsigned v, hipStream_t s);
// ghost selection (incomplete_cells): bit i of sends[p] for neighbour box i
// (6 floats each, [first, nnb)), for original particles p < norig
void launch_ghost_sends(const float4 *pts, const int4 *tets, int ntets, int norig,
                        const float *bmin, const float *bmax, const float *boxes, int nnb,
                        int first, unsigned long long *sends, hipStream_t s);
// xyz triples -> float4 (xyz, 0)
void launch_pack_points(const float *xyz, int np, float4 *pts, hipStream_t s);
// index-range check of a block in device memory: sets `bit` in *err on a violation
void launch_check_block(const int4 *tets, int nt, const int *v2t, int np, unsigned *err,
                        unsigned bit, hipStream_t s);
// zero every block's density and count grids; circumcenters of every block
void launch_block_prep(const BlockIn *ins, const CellOut *outs, int nb, int max_tets,
                       unsigned long long max_grid, bool circumcenters, hipStream_t s);
void launch_cells_ref(const BlockIn &in, const GridP &gp, const CellOut &o,
                      int want_density, hipStream_t s);
void launch_cells_ref_slow(const BlockIn &in, const GridP &gp, const CellOut &o,
                           int want_density, hipStream_t s);
// several cells per wave in the geometry phase (cells_mc.hip); variant 1..2
// picks the verdict masks; returns -1 for an unknown variant.  gp is the
// block's grid descriptor in device memory (read only by the scan, so that
// its fields take no registers in the geometry phase; may be null without
// density)
int launch_cells_mc(const BlockIn &in, const GridP *gp, const CellOut &o, int want_density,
                    int variant, hipStream_t s);
// large-capacity pass over the cells every block's fast pass deferred
void launch_cells_ref_slow_batched(const BlockIn *ins, const Router *rt, int want_density,
                                   hipStream_t s);
// the two-cell kernel over all blocks in one persistent launch (the deferred
// cells then need launch_cells_ref_batched's large-capacity pass)
// the batched launch over an item list (two consecutive cells of one block
// per item), one workgroup per item, no work queue
void launch_cells_mc_items(const BlockIn *ins, const GridP *gps, const CellOut *outs,
                           const Router *rt, const Chunk *items, const XcdQueues &xr,
                           int want_density, hipStream_t s);
void launch_cells_mc_batched(const BlockIn *ins, const GridP *gps, const CellOut *outs,
                             const Router *rt, const Chunk *chunks, const XcdQueues &xq,
                             unsigned *queue, int waves, int want_density, hipStream_t s);
void launch_cic(const BlockIn &in, const GridP &gp, const CellOut &o, const Router *rt_dev,
                int self, hipStream_t s);
// spec: each star-search candidate loads its tet's record before the table
// decides whether it is new (the loads overlap the LDS work)
void launch_cells_ref_batched(const BlockIn *ins, const GridP *gps, const CellOut *outs,
                              const Router *rt, const Chunk *chunks, const XcdQueues &xq,
                              unsigned *queue, int waves, int want_density, hipStream_t s,
                              bool spec);
// ordered fold of shared grid points and per-block statistics for nb blocks
// at once (outs/gps: device arrays; the capacities size the grids)
// nb <= 128; the sums are the list capacities of all blocks
void launch_fold(const CellOut *outs, const GridP *gps, int nb, unsigned max_ovf_cap,
                 const unsigned max_fold_cap[3], unsigned long long max_grid,
                 unsigned long long sum_ovf_cap, unsigned long long sum_mid_cap, hipStream_t s);
// part: nb * 1024 doubles of scratch for the fixed-order grid sums
void launch_stats(const CellOut *outs, double *part, int nb, int max_cells,
                  unsigned long long max_grid, hipStream_t s);

}  // namespace t2
