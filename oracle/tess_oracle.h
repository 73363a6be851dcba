/*
 * tess_oracle.h -- CPU oracle for the tess2 post-tessellation path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * algorithm (tet circumcenters, Voronoi cell volumes, the Voronoi density
 * estimator and its grid deposit).  It is the checker for the HIP library:
 * only the tests, the smoke check and the CPU-baseline leg of the benchmark
 * may load it.  The product path never calls into it.
 *
 * Every function names the reference file:line it restates.  Arithmetic
 * follows the reference's evaluation order operand by operand in IEEE
 * float32 (no contraction, correctly rounded / and sqrt), with float<->double
 * promotions exactly where the C++ source has them.  Build with
 * -ffp-contract=off and without -ffast-math.
 */
#ifndef TESS_ORACLE_H
#define TESS_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* include/tess/tet.h:4-7 -- tets[i] lies opposite verts[i]; -1 = hull */
typedef struct o_tet_t {
    int verts[4];
    int tets[4];
} o_tet_t;

/* One block: the dblock_t fields the path reads
 * (include/tess/delaunay.h:38-63) plus DBlock::bounds
 * (include/tess/delaunay.hpp:18-23). */
typedef struct o_block_t {
    int num_orig_particles;
    int num_particles;
    const float *particles;     /* xyz, 3*num_particles, originals first */
    int num_tets;
    const o_tet_t *tets;
    const int *vert_to_tet;
    float bounds_min[3];
    float bounds_max[3];
} o_block_t;

/* args_t, include/tess/dense.hpp:48-61 */
typedef struct o_args_t {
    int project;
    float proj_plane[3];
    float mass;
    float data_mins[3];
    float data_maxs[3];
    float grid_phys_mins[3];
    float grid_step_size[3];
    float eps;
    int glo_num_idx[3];
} o_args_t;

/* per-cell outcome of the density estimator (shared coding with the GPU) */
enum {
    O_CELL_SCAN = 0,        /* mass spread over scanned interior points   */
    O_CELL_CIC = 1,         /* scan found nothing: CIC of the site        */
    O_CELL_INCOMPLETE = 2,  /* vert_to_tet == -1 or infinite cell         */
    O_CELL_OUTSIDE = 3,     /* bbox leaves the data bounds                */
    O_CELL_DEGENERATE = 4   /* NaN in the cell bbox (reference: UB)       */
};

/* ---- tet geometry (src/tet.cpp) ---- */
void o_circumcenter(float *center, const o_tet_t *tet, const float *particles);
void o_fill_circumcenters(const o_tet_t *tets, int num_tets,
                          const float *particles, float *out);
int o_neighbor_edges(int v, const o_tet_t *tets, int t, int *nbr_u, int *nbr_t,
                     int cap, int *num_nbrs);
int o_complete(int v, const o_tet_t *tets, int num_tets, int t);
int o_fill_edge_link(int *link, int cap, int v, int u, int ut,
                     const o_tet_t *tets);

/* ---- volumes (src/volume.cpp) ---- */
float o_volume(int v, const int *vert_to_tet, const o_tet_t *tets,
               int num_tets, const float *particles, const float *cc);
void o_volumes(const o_block_t *b, const float *cc, float *out);

/* ---- density helpers (src/dense.cpp) ---- */
void o_newell_normal(const float *verts, int num_verts, float *normal);
int o_pt_in_cell(const float *pt, const float *normals, const float *f0,
                 int num_faces, float eps);
void o_idx2phys(const int *idx, float *pos, const float *step,
                const float *gmin);
void o_phys2idx(const float *pos, int *idx, const float *step,
                const float *gmin);
int o_index(const int *bidx, const int *bnum, int project,
            const float *proj_plane);
void o_grid_step_params(int num_given_bounds, const float *given_mins,
                        const float *given_maxs, const float *data_mins,
                        const float *data_maxs, float *grid_phys_mins,
                        float *grid_phys_maxs, float *grid_step_size,
                        const int *glo_num_idx);
void o_block_grid_params(const float *bmin, const float *bmax, int *block_min_idx,
                         int *block_max_idx, int *block_num_idx,
                         const float *grid_phys_mins, const float *grid_step_size,
                         float eps, const float *data_mins,
                         const float *data_maxs, const int *glo_num_idx);
void o_distribute_cic(const float *pt, float scalar, int *idx24, float *mass8,
                      const float *step, const float *gmin, float eps);

/* Face planes and bbox of one cell (CellBounds, src/dense.cpp:657-736).
 * Returns the number of faces, or -1 when cap is too small.  normals and f0
 * receive 3 floats per face; bmin/bmax the cell bbox. */
int o_cell_bounds(const o_block_t *b, int cell, float *bmin, float *bmax,
                  float *normals, float *f0, int cap);

/* The CellInteriorGridPts scan state machine (src/dense.cpp:1475-1701) on a
 * box of cgp[0] x cgp[1] x cgp[2] points whose PtInCell verdicts are given
 * as a mask (1 = inside, x fastest).  Fills border[2*(zi*cgp[1] + yi)] =
 * {min_xi, max_xi} for every line (min > max: no points) and returns the
 * total point count.  o_cell_points runs the same walk on PtInCell. */
int o_scan_mask(const int *cgp, const unsigned char *inside, int *border);

/* Grid points of one cell (IterateCells body up to the deposit,
 * src/dense.cpp:250-268 + CellGridPts :1363-1458).  Writes global grid
 * indices (3 ints each) and per-point masses (double, as grid_pt_t::mass).
 * Returns the number of points (0 when the cell is skipped); *status gets an
 * O_CELL_* code.  cap = capacity in points; returns -1 if exceeded. */
int o_cell_points(const o_block_t *b, const o_args_t *a, int cell, int *idx3,
                  double *mass, int cap, int *status);

/* IterateCells (src/dense.cpp:220-313) for one block.  density holds
 * block_num_idx points (or a 2-D plane when projecting), is updated in place.
 * Optional outputs (may be NULL): per-cell status and point count; off-block
 * points (global idx3 + mass + source cell), up to off_cap.  Returns the
 * number of off-block points produced (may exceed off_cap). */
long o_iterate_cells(const o_block_t *b, const o_args_t *a,
                     const int *block_min_idx, const int *block_num_idx,
                     float *density, int *cell_status, int *cell_npts,
                     int *off_idx3, double *off_mass, int *off_cell,
                     long off_cap, double *tot_mass);

/* FP64 counterpart of o_iterate_cells for error reports: the same cells,
 * point sets and in-block rule, but each scan point's mass is mass/N in
 * double (the reference divides in float, :1692; CIC corners keep the
 * reference's float weights) and density64 (zeroed by the caller)
 * accumulates m/div in double, div the product of the steps in double.
 * Returns the number of off-block points skipped. */
long o_iterate_cells_fp64(const o_block_t *b, const o_args_t *a,
                          const int *block_min_idx, const int *block_num_idx,
                          double *density64);

/* IterateCellsCic (src/dense.cpp:486-562) for one block, same conventions. */
long o_iterate_cells_cic(const o_block_t *b, const o_args_t *a,
                         const int *block_min_idx, const int *block_num_idx,
                         float *density, int *off_idx3, double *off_mass,
                         int *off_cell, long off_cap, double *tot_mass);

/* recvd_pts body (src/dense.cpp:182-200): add received points in order. */
void o_recvd_pts(const int *idx3, const double *mass, long n,
                 const int *block_min_idx, const int *block_num_idx,
                 const o_args_t *a, float *density);

/* ---- ghost selection (src/tess.cpp:492-628) ----
 * incomplete_cells() for one block: for every original particle p, bit i of
 * sends[p] is set when p must go to neighbour i (first <= i < nnb <= 64).
 * boxes: 6 floats per neighbour (min xyz, max xyz), already wrapped; the
 * block's own box is b->bounds_min / bounds_max. */
void o_ghost_sends(const o_block_t *b, int nnb, const float *boxes, int first,
                   unsigned long long *sends);

/* ---- exact Voronoi assignment (FP64 oracle for the exact mode) ----
 * Owner of every grid point of a (nx,ny,nz) grid at pos = idx*step + gmin
 * (float32, as idx2phys): the nearest site among all particles, ties to the
 * lowest particle id, squared distances compared in double.  Brute force
 * over a uniform bucket hash; owner[-1] never happens for a non-empty set. */
void o_exact_owner(const float *particles, int num_particles, const int *gnum,
                   const float *step, const float *gmin, int *owner);

/* Same, for the sub-grid of global indices [off, off + gnum) (a block grid). */
void o_exact_owner_off(const float *particles, int num_particles, const int *off,
                       const int *gnum, const float *step, const float *gmin, int *owner);

/* Exact-Voronoi density of one block (FP64 oracle of the exact mode).
 * A cell deposits when it is complete and the FP64 bounding box of its
 * Voronoi vertices (circumcenters of its star, solved in double; flat tets,
 * whose circumcenter is undefined (NaN), are skipped, and a cell with none
 * defined is degenerate) stays inside the data bounds with the reference's
 * 2*FLT_EPSILON*extent slack.  Every
 * in-block grid point goes to its nearest site (o_exact_owner); a deposited
 * cell with N points gives each mass/N, one with none spreads its mass by
 * the reference's CIC.  Contributions are summed in double in ascending cell
 * order; density = sum / div.  owner (block grid) gets the nearest site of
 * each in-block point (-1 outside the block), status/npts as O_CELL_*. */
void o_exact_dense(const o_block_t *b, const o_args_t *a, const int *block_min_idx,
                   const int *block_num_idx, int *owner, int *status, int *npts,
                   double *density);

#ifdef __cplusplus
}
#endif
#endif
