/*
 * tess2_amd.h -- C ABI of the MI355X (gfx950) implementation of tess2's
 * post-tessellation path: tet circumcenters, Voronoi cell volumes, and the
 * Voronoi density estimator with its deposit onto a regular float32 grid.
 *
 * Plain C types only (pointers + sizes); every entry point returns 0 on
 * success and a negative TESS2_E* code on failure (never aborts), with a
 * message available from tess2_last_error().  Host pointers are the
 * caller's; the library owns its device buffers and its HIP stream.
 *
 * Reference interfaces replaced (paths relative to diatomic/tess2):
 *   tess2_fill_circumcenters  <- fill_circumcenters()  src/volume.cpp:6-11,
 *                                include/tess/volume.h:7-10
 *   tess2_volumes             <- volume() over all original cells,
 *                                src/volume.cpp:13-54, include/tess/volume.h:12-16
 *   tess2_dense               <- dense() + est_dense() + recvd_pts(),
 *                                src/dense.cpp:30-202, include/tess/dense.hpp:75-98
 *   tess2_grid_step_params    <- GridStepParams()  src/dense.cpp:1712-1767
 *   tess2_block_grid_params   <- BlockGridParams() src/dense.cpp:575-648
 *   tess2_write_grid          <- WriteGrid() raw float32 output of one global
 *                                array, src/dense.cpp:751-870
 *   tess2_write_blocks        <- WriteGrid() over blocks (subarray writes,
 *                                ProjectGrid first when projecting),
 *                                src/dense.cpp:751-1023
 *   tess2_project_grid        <- ProjectGrid(), src/dense.cpp:881-1023
 *   tess2_ghost_sends         <- the test of incomplete_cells(),
 *                                src/tess.cpp:492-628
 * The staged (device-resident) calls have no reference counterpart: they
 * split dense() into upload / run / download so that a caller can keep the
 * tessellation resident in HBM across runs.
 */
#ifndef TESS2_AMD_H
#define TESS2_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

/* tet_t, include/tess/tet.h:4-7: tets[i] lies opposite verts[i], -1 = hull */
typedef struct tess2_tet_t {
    int verts[4];
    int tets[4];
} tess2_tet_t;

/* The dblock_t fields the path reads (include/tess/delaunay.h:38-63) plus
 * DBlock::bounds (include/tess/delaunay.hpp:18-23).  Original particles come
 * first, then ghosts. */
typedef struct tess2_block_t {
    int gid;
    int num_orig_particles;
    int num_particles;
    const float *particles;          /* 3*num_particles floats, xyz */
    int num_tets;
    const tess2_tet_t *tets;
    const int *vert_to_tet;          /* num_particles entries, -1 = none */
    float bounds_min[3];
    float bounds_max[3];
    /* optional (NULL when absent), for the exact assignment over several
       blocks: for each ghost (particles num_orig_particles..num_particles-1)
       the gid of the block that holds it as an original, and its index
       there.  No reference counterpart: the reference's ghost exchange
       (src/tess.cpp:492-628) knows this but does not keep it. */
    const int *ghost_gid;
    const int *ghost_cell;
} tess2_block_t;

/* estimator algorithm, enum alg in include/tess/dense.hpp:33-38 */
enum { TESS2_DENSE_TESS = 0, TESS2_DENSE_CIC = 1 };

/* grid-point assignment of the TESS estimator */
enum {
    TESS2_ASSIGN_REFERENCE_SCAN = 0, /* the reference scanline search, bit-exact */
    TESS2_ASSIGN_EXACT_VORONOI = 1   /* nearest site, ties to the lowest id
                                        (several blocks: lowest home key; the
                                        blocks then need ghost_gid/ghost_cell) */
};

/* arithmetic of the EXACT_VORONOI assignment */
enum {
    TESS2_PREC_MIXED = 0,  /* FP32 filters, FP64 fallback near ties */
    TESS2_PREC_FP64 = 1    /* every predicate in FP64               */
};

/* per-cell outcome codes written by tess2_dense_cells() */
enum {
    TESS2_CELL_SCAN = 0,
    TESS2_CELL_CIC = 1,
    TESS2_CELL_INCOMPLETE = 2,
    TESS2_CELL_OUTSIDE = 3,
    TESS2_CELL_DEGENERATE = 4
};

/* error codes */
enum {
    TESS2_OK = 0,
    TESS2_E_ARG = -1,
    TESS2_E_HIP = -2,
    TESS2_E_CAPACITY = -3,   /* a per-cell capacity was exceeded */
    TESS2_E_NOMEM = -4,
    TESS2_E_IO = -5,
    TESS2_E_INTERNAL = -6    /* an internal consistency check failed: outputs are invalid */
};

/* counters of one dense run (the reference prints tot_mass / check_mass /
 * max_dense, src/dense.cpp:25-27,1284-1333) */
typedef struct tess2_stats_t {
    long long cells_scan;
    long long cells_cic;
    long long cells_incomplete;
    long long cells_outside;
    long long cells_degenerate;
    long long points_deposited;      /* in-block deposits                     */
    long long points_offblock;       /* points no block grid holds (dropped)  */
    long long points_routed;         /* points added to another block's grid  */
    long long points_shared;         /* deposits folded by the ordered pass   */
    long long fp64_fallbacks;        /* EXACT_VORONOI mixed: FP64 re-checks   */
    double tot_mass;                 /* sum of deposited point masses         */
    float max_dense;
    float device_ms;                 /* device time of the last run          */
    long long cells_deferred;        /* cells above the one-wave kernel's star or face
                                        capacity, handled by the large-capacity pass */
} tess2_stats_t;

typedef struct tess2_ctx tess2_ctx;

int tess2_create(tess2_ctx **ctx, int device);
int tess2_destroy(tess2_ctx *ctx);
const char *tess2_last_error(const tess2_ctx *ctx);

/* fill_circumcenters(): one circumcenter per tet, float32, bit-exact with
 * src/tet.cpp:37-66.  out holds 3*num_tets floats. */
int tess2_fill_circumcenters(tess2_ctx *ctx, const tess2_tet_t *tets, int num_tets,
                             const float *particles, int num_particles, float *out);

/* volume() for every original cell v < num_orig_particles; -1 for an
 * infinite cell (and for a particle in no tet). */
int tess2_volumes(tess2_ctx *ctx, const tess2_block_t *block, float *out);

/* volume() for every original cell, like tess2_volumes, allocating nothing.
 * No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Every array is in DEVICE memory, owned by the caller:
 * particles (3*num_particles floats, xyz), tets (num_tets tet_t records,
 * 16-byte aligned), vert_to_tet (num_particles), out (num_orig_particles).
 * The library's working memory is one fixed array of 64 MiB per GPU, held
 * from load to process end; a block needing more (about 16 bytes per
 * particle, 16 per tet and 12 per original cell) returns TESS2_E_CAPACITY
 * before any kernel runs, and out-of-range indices return TESS2_E_ARG.
 * Runs on the context's stream and returns when the volumes are written.
 * Every call on one GPU shares that array: calls that overlap on one GPU
 * (two threads, two contexts) race and must be serialised by the caller. */
int tess2_volumes_noalloc(tess2_ctx *ctx, const float *particles, int num_particles,
                          int num_orig_particles, const tess2_tet_t *tets, int num_tets,
                          const int *vert_to_tet, float *out);

/* The test of incomplete_cells() (src/tess.cpp:492-628) for one block of a
 * tessellation round: for every original particle p, bit i of sends[p]
 * (num_orig_particles entries) is set when p must be sent to neighbour i as
 * a ghost: a tet with p has a circumsphere reaching neighbour box i (unless
 * the sphere lies deep inside the block's own bounds_min/bounds_max), or p
 * is on a convex-hull facet whose outside reaches box i (side_of_plane,
 * src/tet.cpp:70-119).  neighbor_boxes: 6 floats per neighbour (min xyz,
 * max xyz), already wrapped for periodic neighbours; at most 64; neighbours
 * below first_neighbor (the reference's last_neighbor) are not tested.
 * The rounds themselves (send, tessellate again, stop when nothing is sent,
 * src/tess.cpp:79-98) are the caller's. */
int tess2_ghost_sends(tess2_ctx *ctx, const tess2_block_t *block, int num_neighbors,
                      const float *neighbor_boxes, int first_neighbor,
                      unsigned long long *sends);

/* GridStepParams() / BlockGridParams() host helpers (same arguments). */
void tess2_grid_step_params(int num_given_bounds, const float *given_mins,
                            const float *given_maxs, const float *data_mins,
                            const float *data_maxs, float *grid_phys_mins,
                            float *grid_phys_maxs, float *grid_step_size,
                            const int *glo_num_idx);
void tess2_block_grid_params(const float *bounds_min, const float *bounds_max,
                             int *block_min_idx, int *block_max_idx,
                             int *block_num_idx, const float *grid_phys_mins,
                             const float *grid_step_size, float eps,
                             const float *data_mins, const float *data_maxs,
                             const int *glo_num_idx);

/* dense(): density of one or more blocks on their block grids.
 * Outputs data_mins/maxs, grid_phys_mins/maxs, grid_step_size as the
 * reference does; density_out[b] must hold the block's grid
 * (block_num_idx product, or x*y when projecting).  assignment /
 * precision select the TESS estimator variant (REFERENCE_SCAN is the
 * reference's algorithm).  stats may be NULL. */
int tess2_dense(tess2_ctx *ctx, int alg, int num_given_bounds,
                const float *given_mins, const float *given_maxs, int project,
                const float *proj_plane, float mass, float *data_mins,
                float *data_maxs, float *grid_phys_mins, float *grid_phys_maxs,
                float *grid_step_size, float eps, const int *glo_num_idx,
                const tess2_block_t *blocks, int nblocks, int assignment,
                int precision, float **density_out, tess2_stats_t *stats);

/* Per-cell detail of the last tess2_dense / tess2_run call for block b:
 * status (TESS2_CELL_*) and point count per original cell. */
int tess2_dense_cells(tess2_ctx *ctx, int b, int *status, int *npts);

/* WriteGrid(): raw float32, native endian, C order (x fastest), no header;
 * the file is truncated first (src/dense.cpp:781-870).  With project, the
 * 2-D [y][x] plane. */
int tess2_write_grid(const char *path, const float *density, const int *glo_num_idx,
                     int project);

/* ProjectGrid() for blocks all held by the caller (one process): in
 * ascending gid order, every block whose grid starts above z index 0 adds
 * its 2-D plane, float += float point by point, into the plane of the block
 * with the same x-y grid start at z index 0 (the root), as the reference
 * does for blocks on one rank (src/dense.cpp:881-1023).  density[i] is block
 * i's [y][x] plane from a projecting tess2_dense; block_min_idx /
 * block_num_idx hold 3 ints per block (tess2_block_grid_params).  The roots'
 * planes are updated in place. */
int tess2_project_grid(int nblocks, const int *gids, float **density,
                       const int *block_min_idx, const int *block_num_idx);

/* WriteGrid() for blocks all held by the caller: the file is truncated,
 * then every block's grid is written at its place in the global C-order
 * array (x fastest).  With project, tess2_project_grid runs first (updating
 * the roots' planes in place, as the reference's WriteGrid does) and only
 * the blocks at z index 0 write their [y][x] planes into the global
 * glo_num_idx[1] x glo_num_idx[0] plane (src/dense.cpp:800-846). */
int tess2_write_blocks(const char *path, int nblocks, const int *gids, float **density,
                       const int *block_min_idx, const int *block_num_idx,
                       const int *glo_num_idx, int project);

/* ---- staged (device-resident) interface ---- */

/* Upload blocks once; later runs read them from HBM. */
int tess2_stage(tess2_ctx *ctx, const tess2_block_t *blocks, int nblocks);

/* Run the whole path on the staged blocks, asynchronously on the context
 * stream: circumcenters -> stars -> volumes -> face planes -> assignment ->
 * deposit.  Arguments as tess2_dense (data bounds are taken from the staged
 * blocks).  Results stay on the device until tess2_fetch. */
int tess2_run(tess2_ctx *ctx, int alg, int num_given_bounds, const float *given_mins,
              const float *given_maxs, int project, const float *proj_plane,
              float mass, float eps, const int *glo_num_idx, int assignment,
              int precision, int want_volumes);

/* Wait for the last run; fill stats (device_ms = event time of the run). */
int tess2_sync(tess2_ctx *ctx, tess2_stats_t *stats);

/* Copy results of block b to the host (NULL pointers are skipped).
 * Volumes exist only after a TESS run with want_volumes (or tess2_volumes);
 * asking for them after any other run returns TESS2_E_ARG. */
int tess2_fetch(tess2_ctx *ctx, int b, float *density, float *volumes,
                float *circumcenters);

/* Nearest-site owner of every grid point of block b from the last
 * EXACT_VORONOI run (-1 outside the block bounds); block grid order.  One
 * staged block: the particle index; several: the home key
 * (rank << 23 | cell, rank = position in ascending gid order). */
int tess2_fetch_owner(tess2_ctx *ctx, int b, int *owner);

/* Average device time per launch (ms) of the dominant kernel over the
 * runs since the last call, measured with HIP events on the library
 * stream, and the number of launches averaged. */
int tess2_kernel_time(tess2_ctx *ctx, float *avg_ms, int *launches);

/* ---- launch options and tuning (no counterpart in the reference: its
 * dense() runs one OpenMP-free loop per block, src/dense.cpp:216-280) ---- */

enum {
    TESS2_OPT_GRAPH = 0,          /* 1: record the run's launches once into a HIP graph and
                                   * replay it while arguments and buffers are unchanged */
    TESS2_OPT_WAVES_PER_CU = 1,   /* persistent waves per CU of the multi-block cell kernel */
    TESS2_OPT_CHUNK_SMALL = 2,    /* cells per work item, blocks with < 20000 cells */
    TESS2_OPT_CHUNK_MEDIUM = 3,   /* cells per work item, blocks with < 100000 cells */
    TESS2_OPT_CHUNK_LARGE = 4,    /* cells per work item, larger blocks */
    TESS2_OPT_CELL_KERNEL = 5     /* cell kernel: -1 (default) chosen per run (one block:
                                   * 1, or 2 when cell boxes are large, 16 or more grid
                                   * points per cell; several blocks: 3); 0 one cell per
                                   * wave; 1 two cells per wave in the star / face / ring
                                   * phase; 2 the same with the verdict masks of boxes up
                                   * to 512 points (several blocks: 1 and 2 both run the
                                   * two-cell persistent kernel with those masks); 3 one
                                   * cell per wave with the star search's record loads
                                   * issued before its table work (several blocks only;
                                   * one block runs 0); 4 several blocks in one launch
                                   * without the work queue, a workgroup per two cells,
                                   * two cells per wave as 2 (one block runs 2);
                                   * with the option at -1, tess2_autotune also times
                                   * 1 and 4 and keeps the fastest */
};

/* Set / read one option.  Chunk sizes 1..4096, waves per CU 1..64, cell
 * kernel -1..4. */
int tess2_set_option(tess2_ctx *ctx, int option, int value);
int tess2_get_option(tess2_ctx *ctx, int option, int *value);

/* Time the last tess2_run's arguments under each candidate work split of the
 * multi-block cell kernel (chunk sizes x persistent waves per CU), and, when
 * TESS2_OPT_CELL_KERNEL is -1, the two-cell kernels (cell kernels 1 and 4), `reps`
 * runs each, keep the fastest, and leave it set.  Results of the staged
 * blocks are those of the last timed run (identical for every candidate).
 * best_ms (may be NULL) receives the fastest median run time. */
int tess2_autotune(tess2_ctx *ctx, int reps, float *best_ms);

#ifdef __cplusplus
}
#endif
#endif
