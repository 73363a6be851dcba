/*
 * tess_oracle.c -- CPU oracle for the tess2 post-tessellation path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library, loaded by the
 * tests, the smoke check and the benchmark's CPU-baseline leg, never by the
 * product path.
 *
 * A clean-room restatement of the reference's algorithm, following
 * src/tet.cpp, src/volume.cpp and src/dense.cpp operation by operation.  Each
 * function cites the reference lines it follows.  Float arithmetic is IEEE
 * binary32 in the reference's evaluation order; where the C++ source mixes
 * float and double (a double literal, a double-typed struct field, a call to
 * the double sqrt) the promotion is written out explicitly here.
 *
 * Parity status of each part (what pins it) is listed in the test files
 * tests/test_oracle_*.py next to the checks themselves.
 */
#include "tess_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* small growable int buffer (scratch for the BFS / edge links)              */
/* ------------------------------------------------------------------------ */
typedef struct {
    int *p;
    int n, cap;
} ibuf;

static void ib_push(ibuf *b, int x)
{
    if (b->n == b->cap) {
        b->cap = b->cap ? 2 * b->cap : 64;
        b->p = (int *)realloc(b->p, sizeof(int) * (size_t)b->cap);
    }
    b->p[b->n++] = x;
}

static int ib_has(const ibuf *b, int x)
{
    for (int i = 0; i < b->n; ++i)
        if (b->p[i] == x)
            return 1;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* tet geometry                                                             */
/* ------------------------------------------------------------------------ */

/* norm(): |x|^2, src/tet.cpp:122-128 (res starts at 0 and accumulates) */
static float o_norm(const float *x)
{
    float r = 0.0f;
    for (int i = 0; i < 3; ++i)
        r += x[i] * x[i];
    return r;
}

/* cross(): src/tet.cpp:131-136 */
static void o_cross(float *r, const float *u, const float *v)
{
    r[0] = u[1] * v[2] - u[2] * v[1];
    r[1] = u[2] * v[0] - u[0] * v[2];
    r[2] = u[0] * v[1] - u[1] * v[0];
}

/* determinant(): six triple products, left to right, src/tet.cpp:139-143 */
static float o_det(const float *t, const float *u, const float *v)
{
    return t[0] * u[1] * v[2] + u[0] * v[1] * t[2] + v[0] * t[1] * u[2] -
           v[0] * u[1] * t[2] - u[0] * t[1] * v[2] - t[0] * v[1] * u[2];
}

/* distance(): sqrt of the float sum; the float result of the double sqrt
 * equals the correctly rounded float sqrt, src/tet.cpp:146-153 */
static float o_distance(const float *u, const float *v)
{
    float n = 0.0f;
    for (int i = 0; i < 3; ++i)
        n += (u[i] - v[i]) * (u[i] - v[i]);
    return sqrtf(n);
}

/* circumcenter(): src/tet.cpp:37-66 */
void o_circumcenter(float *center, const o_tet_t *tet, const float *particles)
{
    const float *a = &particles[3 * tet->verts[0]];
    const float *b = &particles[3 * tet->verts[1]];
    const float *c = &particles[3 * tet->verts[2]];
    const float *d = &particles[3 * tet->verts[3]];
    float t[3], u[3], v[3];
    for (int i = 0; i < 3; ++i) {
        t[i] = a[i] - d[i];
        u[i] = b[i] - d[i];
        v[i] = c[i] - d[i];
    }
    float nt = o_norm(t), nu = o_norm(u), nv = o_norm(v);
    float den = 2.0f * o_det(t, u, v);
    float uv[3], vt[3], tu[3];
    o_cross(uv, u, v);
    o_cross(vt, v, t);
    o_cross(tu, t, u);
    for (int i = 0; i < 3; ++i)
        center[i] = d[i] + (nt * uv[i] + nu * vt[i] + nv * tu[i]) / den;
}

/* fill_circumcenters(): src/volume.cpp:6-11 */
void o_fill_circumcenters(const o_tet_t *tets, int num_tets,
                          const float *particles, float *out)
{
    for (int i = 0; i < num_tets; ++i)
        o_circumcenter(&out[3 * i], &tets[i], particles);
}

/* BFS over the star of v from tet t, src/tet.cpp:228-270.  The queue may
 * hold a tet several times; a tet is expanded on its first pop.  Fills
 * (u, ut) pairs in first-seen order.  Returns finite (1/0). */
static int star_bfs(int v, const o_tet_t *tets, int t, ibuf *q, ibuf *vis,
                    ibuf *vverts, ibuf *nu, ibuf *nt)
{
    int finite = 1;
    q->n = vis->n = vverts->n = 0;
    if (nu) nu->n = 0;
    if (nt) nt->n = 0;
    ib_push(q, t);
    for (int head = 0; head < q->n; ++head) {
        int cur = q->p[head];
        if (ib_has(vis, cur))
            continue;
        ib_push(vis, cur);
        for (int i = 0; i < 4; ++i) {
            int u = tets[cur].verts[i];
            if (u != v) {
                if (!ib_has(vverts, u)) {
                    ib_push(vverts, u);
                    if (nu) ib_push(nu, u);
                    if (nt) ib_push(nt, cur);
                }
                int next = tets[cur].tets[i];
                if (next == -1)
                    finite = 0;
                else
                    ib_push(q, next);
            }
        }
    }
    return finite;
}

/* scratch buffers shared by the per-cell routines (single-threaded oracle) */
static ibuf g_q, g_vis, g_vv, g_nu, g_nt, g_link;

int o_neighbor_edges(int v, const o_tet_t *tets, int t, int *nbr_u, int *nbr_t,
                     int cap, int *num_nbrs)
{
    int finite = star_bfs(v, tets, t, &g_q, &g_vis, &g_vv, &g_nu, &g_nt);
    *num_nbrs = g_nu.n;
    if (g_nu.n > cap)
        return -1;
    for (int i = 0; i < g_nu.n; ++i) {
        nbr_u[i] = g_nu.p[i];
        nbr_t[i] = g_nt.p[i];
    }
    return finite;
}

/* complete(): same BFS, early exit on a hull neighbour, src/tet.cpp:337-378 */
int o_complete(int v, const o_tet_t *tets, int num_tets, int t)
{
    (void)num_tets;
    g_q.n = g_vis.n = 0;
    ib_push(&g_q, t);
    for (int head = 0; head < g_q.n; ++head) {
        int cur = g_q.p[head];
        if (ib_has(&g_vis, cur))
            continue;
        ib_push(&g_vis, cur);
        for (int i = 0; i < 4; ++i) {
            int u = tets[cur].verts[i];
            if (u != v) {
                int next = tets[cur].tets[i];
                if (next == -1)
                    return 0;
                ib_push(&g_q, next);
            }
        }
    }
    return 1;
}

/* circulate_start(): first slot holding neither edge vertex,
 * src/tet.cpp:164-173 */
static int circ_start(const o_tet_t *tets, int t, int x, int y)
{
    for (int i = 0; i < 4; ++i) {
        int w = tets[t].verts[i];
        if (w != x && w != y)
            return i;
    }
    return -1;
}

/* circulate_next(): cross the face opposite slot wi, src/tet.cpp:185-215 */
static void circ_next(int *next_t, int *next_v, const o_tet_t *tets, int t,
                      int wi, int x, int y)
{
    const o_tet_t *cur = tets + t;
    int nv = -1;
    for (int i = 0; i < 4; ++i) {
        int w = cur->verts[i];
        if (i != wi && w != x && w != y) {
            nv = w;
            break;
        }
    }
    *next_t = cur->tets[wi];
    *next_v = -1;
    if (*next_t < 0)
        return;
    const o_tet_t *nx = tets + *next_t;
    for (int i = 0; i < 4; ++i)
        if (nx->verts[i] == nv) {
            *next_v = i;
            return;
        }
}

/* fill_edge_link(): ring of tets around edge (v,u) from ut,
 * src/tet.cpp:389-409 */
static void edge_link(ibuf *link, int v, int u, int ut, const o_tet_t *tets)
{
    link->n = 0;
    int wi = circ_start(tets, ut, v, u);
    int t = ut;
    for (;;) {
        ib_push(link, t);
        if (wi < 0)     /* malformed ring (reference: assert); stop here */
            break;
        int nt, nwi;
        circ_next(&nt, &nwi, tets, t, wi, v, u);
        if (nt == ut || nt == -1)
            break;
        t = nt;
        wi = nwi;
    }
}

int o_fill_edge_link(int *link, int cap, int v, int u, int ut,
                     const o_tet_t *tets)
{
    edge_link(&g_link, v, u, ut, tets);
    if (g_link.n > cap)
        return -1;
    memcpy(link, g_link.p, sizeof(int) * (size_t)g_link.n);
    return g_link.n;
}

/* ------------------------------------------------------------------------ */
/* volume(): src/volume.cpp:13-54                                            */
/* ------------------------------------------------------------------------ */
static float volume_impl(int v, int vt, const o_tet_t *tets,
                         const float *particles, const float *cc)
{
    int finite = star_bfs(v, tets, vt, &g_q, &g_vis, &g_vv, &g_nu, &g_nt);
    if (!finite)
        return -1.0f;
    float vol = 0.0f;
    for (int k = 0; k < g_nu.n; ++k) {
        int u = g_nu.p[k];
        int ut = g_nt.p[k];
        edge_link(&g_link, v, u, ut, tets);
        float area = 0.0f;
        int a = g_link.p[0];
        for (int i = 1; i < g_link.n - 1; ++i) {
            int bb = g_link.p[i];
            int c = g_link.p[i + 1];
            float ab[3], ac[3], cp[3];
            for (int j = 0; j < 3; ++j) {
                ab[j] = cc[3 * bb + j] - cc[3 * a + j];
                ac[j] = cc[3 * c + j] - cc[3 * a + j];
            }
            o_cross(cp, ab, ac);
            /* area += sqrt(norm(cp))/2 with the double sqrt: the float sum is
             * promoted, the double result rounded back (volume.cpp:45) */
            area = (float)((double)area + sqrt((double)o_norm(cp)) / 2.0);
        }
        float dist = o_distance(&particles[3 * u], &particles[3 * v]);
        vol += area * dist / 6.0f;
    }
    return vol;
}

float o_volume(int v, const int *vert_to_tet, const o_tet_t *tets,
               int num_tets, const float *particles, const float *cc)
{
    (void)num_tets;
    int vt = vert_to_tet[v];
    if (vt < 0)     /* reference: undefined (tets[-1]); defined here as -1 */
        return -1.0f;
    return volume_impl(v, vt, tets, particles, cc);
}

void o_volumes(const o_block_t *b, const float *cc, float *out)
{
    for (int v = 0; v < b->num_orig_particles; ++v)
        out[v] = o_volume(v, b->vert_to_tet, b->tets, b->num_tets,
                          b->particles, cc);
}

/* ------------------------------------------------------------------------ */
/* density helpers                                                          */
/* ------------------------------------------------------------------------ */

/* NewellNormal(): src/dense.cpp:1131-1162 */
void o_newell_normal(const float *verts, int num_verts, float *normal)
{
    normal[0] = 0.0f;
    normal[1] = 0.0f;
    normal[2] = 0.0f;
    for (int i = 0; i < num_verts; i++) {
        int cur = i;
        int next = (i + 1) % num_verts;
        normal[0] += (verts[3 * cur + 1] - verts[3 * next + 1]) *
                     (verts[3 * cur + 2] + verts[3 * next + 2]);
        normal[1] += (verts[3 * cur + 2] - verts[3 * next + 2]) *
                     (verts[3 * cur + 0] + verts[3 * next + 0]);
        normal[2] += (verts[3 * cur + 0] - verts[3 * next + 0]) *
                     (verts[3 * cur + 1] + verts[3 * next + 1]);
    }
    float mag = sqrtf(normal[0] * normal[0] + normal[1] * normal[1] +
                      normal[2] * normal[2]);
    normal[0] /= mag;
    normal[1] /= mag;
    normal[2] /= mag;
    normal[0] = -normal[0];
    normal[1] = -normal[1];
    normal[2] = -normal[2];
}

/* PtInCell(): all faces whose |distance| exceeds eps must agree in sign,
 * src/dense.cpp:1172-1203.  Only the first vertex of each face is read. */
int o_pt_in_cell(const float *pt, const float *normals, const float *f0,
                 int num_faces, float eps)
{
    int old_sign = 0;
    for (int k = 0; k < num_faces; k++) {
        const float *n = &normals[3 * k];
        const float *f = &f0[3 * k];
        float dist = n[0] * (pt[0] - f[0]) + n[1] * (pt[1] - f[1]) +
                     n[2] * (pt[2] - f[2]);
        if (fabsf(dist) > eps) {
            int sign = ((double)dist >= 0.0 ? 1 : -1);
            if (old_sign == 0)
                old_sign = sign;
            if (old_sign != sign)
                return 0;
        }
    }
    return 1;
}

/* idx2phys(): src/dense.cpp:1098-1106 */
void o_idx2phys(const int *idx, float *pos, const float *step, const float *gmin)
{
    for (int i = 0; i < 3; ++i)
        pos[i] = (float)idx[i] * step[i] + gmin[i];
}

/* phys2idx(): truncating conversion, src/dense.cpp:1117-1125 */
void o_phys2idx(const float *pos, int *idx, const float *step, const float *gmin)
{
    for (int i = 0; i < 3; ++i)
        idx[i] = (int)((pos[i] - gmin[i]) / step[i]);
}

/* index(): src/dense.cpp:1047-1090 (projection variant for the x-y plane) */
int o_index(const int *bidx, const int *bnum, int project,
            const float *proj_plane)
{
    int idx[3];
    if (project) {
        float pl = (float)bidx[0] * proj_plane[0] + (float)bidx[1] * proj_plane[1] +
                   (float)bidx[2] * proj_plane[2];
        float sh[3] = {pl * proj_plane[0], pl * proj_plane[1], pl * proj_plane[2]};
        for (int i = 0; i < 3; ++i)
            idx[i] = (int)((float)bidx[i] - sh[i]);
    } else {
        idx[0] = bidx[0];
        idx[1] = bidx[1];
        idx[2] = bidx[2];
    }
    return idx[2] * bnum[1] * bnum[0] + idx[1] * bnum[0] + idx[0];
}

/* GridStepParams(): src/dense.cpp:1712-1767.  The half-padding is computed
 * with the double literal 2.0, so the padding step runs in double. */
void o_grid_step_params(int num_given_bounds, const float *given_mins,
                        const float *given_maxs, const float *data_mins,
                        const float *data_maxs, float *grid_phys_mins,
                        float *grid_phys_maxs, float *grid_step_size,
                        const int *glo_num_idx)
{
    float mds = data_maxs[0] - data_mins[0];
    if (data_maxs[1] - data_mins[1] > mds)
        mds = data_maxs[1] - data_mins[1];
    if (data_maxs[2] - data_mins[2] > mds)
        mds = data_maxs[2] - data_mins[2];
    for (int i = 0; i < 3; ++i) {
        float pad = mds - (data_maxs[i] - data_mins[i]);
        grid_phys_mins[i] = (float)((double)data_mins[i] - (double)pad / 2.0);
        grid_phys_maxs[i] = (float)((double)data_maxs[i] + (double)pad / 2.0);
    }
    for (int i = 0; i < 3; ++i)
        if (num_given_bounds >= i + 1) {
            grid_phys_mins[i] = given_mins[i];
            grid_phys_maxs[i] = given_maxs[i];
        }
    for (int i = 0; i < 3; ++i)
        grid_step_size[i] =
            (grid_phys_maxs[i] - grid_phys_mins[i]) / (float)(glo_num_idx[i] - 1);
}

/* BlockGridParams(): src/dense.cpp:575-648 (asserts omitted: the reference
 * is built as Release, CMakeLists.txt:18-22). */
void o_block_grid_params(const float *bmin, const float *bmax, int *block_min_idx,
                         int *block_max_idx, int *block_num_idx,
                         const float *gmin, const float *step, float eps,
                         const float *data_mins, const float *data_maxs,
                         const int *glo_num_idx)
{
    float pos[3];
    o_phys2idx(bmin, block_min_idx, step, gmin);
    o_idx2phys(block_min_idx, pos, step, gmin);
    for (int i = 0; i < 3; ++i)
        if (pos[i] < bmin[i])
            block_min_idx[i]++;
    o_phys2idx(bmax, block_max_idx, step, gmin);
    o_idx2phys(block_max_idx, pos, step, gmin);
    for (int i = 0; i < 3; ++i)
        if (pos[i] + step[i] <= bmax[i])
            block_max_idx[i]++;
    /* duplicate at the shared max face (note: uses data_mins, :615-626) */
    for (int i = 0; i < 3; ++i)
        if (fabsf(data_mins[i] + (float)block_max_idx[i] * step[i] - bmax[i]) < eps &&
            fabsf(bmax[i] - data_maxs[i]) > step[i])
            block_max_idx[i]--;
    for (int i = 0; i < 3; ++i)
        if (fabsf(bmin[i] - data_mins[i]) < step[i])
            block_min_idx[i] = 0;
    for (int i = 0; i < 3; ++i)
        if (fabsf(bmax[i] - data_maxs[i]) < step[i])
            block_max_idx[i] = glo_num_idx[i] - 1;
    for (int i = 0; i < 3; ++i)
        block_num_idx[i] = block_max_idx[i] - block_min_idx[i] + 1;
}

/* DistributeScalarCIC(): src/dense.cpp:1787-1872.  Weight of a corner is
 * v0/vol of the box between the (nudged) point and that corner, normalised;
 * corners in z, y, x order with x fastest. */
void o_distribute_cic(const float *pt, float scalar, int *idx24, float *mass8,
                      const float *step, const float *gmin, float eps)
{
    int mn[3];
    o_phys2idx(pt, mn, step, gmin);
    float tot = 0.0f, v0 = 0.0f, w[8];
    int n = 0, ijk[3];
    for (ijk[2] = mn[2]; ijk[2] <= mn[2] + 1; ijk[2]++)
        for (ijk[1] = mn[1]; ijk[1] <= mn[1] + 1; ijk[1]++)
            for (ijk[0] = mn[0]; ijk[0] <= mn[0] + 1; ijk[0]++) {
                idx24[3 * n + 0] = ijk[0];
                idx24[3 * n + 1] = ijk[1];
                idx24[3 * n + 2] = ijk[2];
                float gp[3], p[3] = {pt[0], pt[1], pt[2]};
                o_idx2phys(ijk, gp, step, gmin);
                for (int i = 0; i < 3; ++i)
                    if (fabsf(p[i] - gp[i]) < eps)
                        p[i] += 2.0f * eps;
                float vol = fabsf((gp[0] - p[0]) * (gp[1] - p[1]) * (gp[2] - p[2]));
                if ((double)v0 == 0.0)
                    v0 = vol;
                float wv = v0 / vol;
                w[n] = wv;
                tot += wv;
                n++;
            }
    for (int i = 0; i < 8; ++i) {
        w[i] /= tot;
        mass8[i] = w[i] * scalar;
    }
}

/* ------------------------------------------------------------------------ */
/* CellBounds(): src/dense.cpp:657-736                                       */
/* ------------------------------------------------------------------------ */
static float *g_fv;     /* ring vertex scratch */
static int g_fv_cap;

int o_cell_bounds(const o_block_t *b, int cell, float *cmin, float *cmax,
                  float *normals, float *f0, int cap)
{
    const o_tet_t *tets = b->tets;
    int t = b->vert_to_tet[cell];
    star_bfs(cell, tets, t, &g_q, &g_vis, &g_vv, &g_nu, &g_nt);
    int nf = g_nu.n;
    if (nf > cap)
        return -1;
    for (int k = 0; k < nf; k++) {
        int u = g_nu.p[k];
        int ut = g_nt.p[k];
        edge_link(&g_link, cell, u, ut, tets);
        int L = g_link.n;
        if (3 * L > g_fv_cap) {
            g_fv_cap = 3 * L + 64;
            g_fv = (float *)realloc(g_fv, sizeof(float) * (size_t)g_fv_cap);
        }
        for (int l = 0; l < L; l++) {
            float vv[3];
            o_circumcenter(vv, &tets[g_link.p[l]], b->particles);
            g_fv[3 * l + 0] = vv[0];
            g_fv[3 * l + 1] = vv[1];
            g_fv[3 * l + 2] = vv[2];
            for (int i = 0; i < 3; ++i) {
                if ((k == 0 && l == 0) || vv[i] < cmin[i])
                    cmin[i] = vv[i];
                if ((k == 0 && l == 0) || vv[i] > cmax[i])
                    cmax[i] = vv[i];
            }
        }
        float n[3];
        o_newell_normal(g_fv, L, n);
        const float *site = &b->particles[3 * cell];
        float w[3] = {g_fv[0] - site[0], g_fv[1] - site[1], g_fv[2] - site[2]};
        if ((double)(w[0] * n[0] + w[1] * n[1] + w[2] * n[2]) < 0.0) {
            n[0] = -n[0];
            n[1] = -n[1];
            n[2] = -n[2];
        }
        normals[3 * k + 0] = n[0];
        normals[3 * k + 1] = n[1];
        normals[3 * k + 2] = n[2];
        f0[3 * k + 0] = g_fv[0];
        f0[3 * k + 1] = g_fv[1];
        f0[3 * k + 2] = g_fv[2];
    }
    return nf;
}

/* ------------------------------------------------------------------------ */
/* CellInteriorGridPts(): the scanline search, src/dense.cpp:1475-1701.      */
/* Every PtInCell probe of the search is at a grid point cmin_pos + i*step   */
/* of the cell's box, so the search is a function of the box's verdicts by   */
/* index; scan_walk runs it on any verdict source and fills the per-line     */
/* runs border[2*(zi*ny + yi)] = {min_xi, max_xi} (min > max: empty line).   */
/* ------------------------------------------------------------------------ */
typedef int (*scan_verdict_fn)(const void *ctx, int xi, int yi, int zi);

static int scan_walk(const int *cgp, scan_verdict_fn inside, const void *ctx,
                     int *border)
{
    int tot = 0;
    int x_left = cgp[0] / 2, x_right = cgp[0] / 2;
    int y_start = 0;
    int x_in_left, x_in_right, min_xi, max_xi;
    int border_found, z_step_done;
    int xi, yi, zi, yj, first_x = 0;

    for (zi = 0; zi < cgp[2]; zi++) {
        border_found = 0;
        z_step_done = 0;
        for (yj = 0; yj < y_start; yj++) {
            border[2 * (zi * cgp[1] + yj)] = 1;
            border[2 * (zi * cgp[1] + yj) + 1] = 0;
        }
        for (yi = y_start; yi < cgp[1]; yi++) {
            /* both stepping flags come from the point at x_left (:1530-1542) */
            x_in_left = x_in_right = inside(ctx, x_left, yi, zi) ? 1 : 0;
            min_xi = cgp[0] - 1;
            max_xi = 0;

            /* x step left (:1547-1582) */
            for (xi = x_left; xi >= 0 && xi < cgp[0];) {
                if (inside(ctx, xi, yi, zi)) {
                    if (xi < min_xi) min_xi = xi;
                    if (xi > max_xi) max_xi = xi;
                    if (x_in_left) {
                        xi--;
                    } else {
                        x_left = xi;
                        break;
                    }
                } else {
                    if (!x_in_left) {
                        xi++;
                    } else {
                        x_left = xi;
                        break;
                    }
                }
            }
            /* x step right (:1585-1620) */
            for (xi = x_right; xi >= 0 && xi < cgp[0];) {
                if (inside(ctx, xi, yi, zi)) {
                    if (xi < min_xi) min_xi = xi;
                    if (xi > max_xi) max_xi = xi;
                    if (x_in_right) {
                        xi++;
                    } else {
                        x_right = xi;
                        break;
                    }
                } else {
                    if (!x_in_right) {
                        xi--;
                    } else {
                        x_right = xi;
                        break;
                    }
                }
            }
            border[2 * (zi * cgp[1] + yi)] = min_xi;
            border[2 * (zi * cgp[1] + yi) + 1] = max_xi;

            if (min_xi <= max_xi) {
                tot += (max_xi - min_xi + 1);
                if (yi == y_start)
                    first_x = (min_xi + max_xi) / 2;
            }
            /* first_y is y_start unless this line is the z step's first
             * with points (:1638-1644) */
            int first_y = y_start;
            if (min_xi <= max_xi && !border_found) {
                first_y = yi;
                border_found = 1;
            }
            if (min_xi > max_xi && border_found)
                z_step_done = 1;

            /* loosen y_start for the next z step: probe down from first_y
             * at x = first_x in the next slice (:1649-1661) */
            if ((yi == cgp[1] - 1 || z_step_done) && zi + 1 < cgp[2]) {
                for (yj = first_y; yj > 0; yj--)
                    if (!inside(ctx, first_x, yj, zi + 1))
                        break;
                y_start = yj;
            }
            if (z_step_done) {
                for (yj = yi + 1; yj < cgp[1]; yj++) {
                    border[2 * (zi * cgp[1] + yj)] = 1;
                    border[2 * (zi * cgp[1] + yj) + 1] = 0;
                }
                break;
            }
        }
    }
    return tot;
}

/* PtInCell verdicts of a cell's box (the probes of :1525-1657) */
typedef struct {
    const float *cmin_pos, *step, *normals, *f0;
    int nf;
    float eps;
} cell_box_t;

static int cell_box_inside(const void *ctx, int xi, int yi, int zi)
{
    const cell_box_t *c = (const cell_box_t *)ctx;
    float gp[3];
    gp[0] = c->cmin_pos[0] + (float)xi * c->step[0];
    gp[1] = c->cmin_pos[1] + (float)yi * c->step[1];
    gp[2] = c->cmin_pos[2] + (float)zi * c->step[2];
    return o_pt_in_cell(gp, c->normals, c->f0, c->nf, c->eps);
}

/* verdicts given as a mask, x fastest */
typedef struct {
    const unsigned char *mask;
    int nx, ny;
} mask_box_t;

static int mask_box_inside(const void *ctx, int xi, int yi, int zi)
{
    const mask_box_t *m = (const mask_box_t *)ctx;
    return m->mask[((long)zi * m->ny + yi) * m->nx + xi] != 0;
}

static int *g_border;
static long g_border_cap;

static int interior_scan(const int *cgp, const float *cmin_pos,
                         const float *normals, const float *f0, int nf,
                         const float *step, float eps)
{
    long need = 2L * cgp[1] * cgp[2];
    if (need > g_border_cap) {
        g_border_cap = need + 64;
        g_border = (int *)realloc(g_border, sizeof(int) * (size_t)g_border_cap);
    }
    cell_box_t c = {cmin_pos, step, normals, f0, nf, eps};
    return scan_walk(cgp, cell_box_inside, &c, g_border);
}

int o_scan_mask(const int *cgp, const unsigned char *inside, int *border)
{
    mask_box_t m = {inside, cgp[0], cgp[1]};
    return scan_walk(cgp, mask_box_inside, &m, border);
}

/* ------------------------------------------------------------------------ */
/* per-cell point list: IterateCells body + CellGridPts                     */
/* ------------------------------------------------------------------------ */
static float *g_norm, *g_f0;
static int g_face_cap;

static int ensure_faces(int nf)
{
    if (nf > g_face_cap) {
        g_face_cap = nf + 64;
        g_norm = (float *)realloc(g_norm, sizeof(float) * 3 * (size_t)g_face_cap);
        g_f0 = (float *)realloc(g_f0, sizeof(float) * 3 * (size_t)g_face_cap);
    }
    return g_face_cap;
}

/* Computes the cell's points into (idx3, mass); returns count or -1 when
 * cap is exceeded.  Mirrors src/dense.cpp:250-268 then :1381-1457. */
int o_cell_points(const o_block_t *b, const o_args_t *a, int cell, int *idx3,
                  double *mass, int cap, int *status)
{
    int vt = b->vert_to_tet[cell];
    if (vt == -1 || !o_complete(cell, b->tets, b->num_tets, vt)) {
        *status = O_CELL_INCOMPLETE;
        return 0;
    }
    /* CellBounds: the face count is the BFS neighbour count */
    int nf;
    for (;;) {
        float cmin[3], cmax[3];
        nf = o_cell_bounds(b, cell, cmin, cmax, g_norm, g_f0, g_face_cap);
        if (nf < 0) {
            ensure_faces(g_nu.n);
            continue;
        }
        /* CellGridPts: data-bounds filter (:1381-1392) */
        float epsv[3];
        for (int i = 0; i < 3; ++i)
            epsv[i] = (a->data_maxs[i] - a->data_mins[i]) * 2.0f * FLT_EPSILON;
        if (cmin[0] < a->data_mins[0] - epsv[0] || cmin[1] < a->data_mins[1] - epsv[1] ||
            cmin[2] < a->data_mins[2] - epsv[2] || cmax[0] > a->data_maxs[0] + epsv[0] ||
            cmax[1] > a->data_maxs[1] + epsv[1] || cmax[2] > a->data_maxs[2] + epsv[2]) {
            *status = O_CELL_OUTSIDE;
            return 0;
        }
        if (isnan(cmin[0]) || isnan(cmin[1]) || isnan(cmin[2]) || isnan(cmax[0]) ||
            isnan(cmax[1]) || isnan(cmax[2])) {
            /* the reference's phys2idx of NaN is undefined; skip the cell */
            *status = O_CELL_DEGENERATE;
            return 0;
        }
        int cmi[3], cma[3], cgp[3];
        float cminpos[3];
        o_phys2idx(cmin, cmi, a->grid_step_size, a->grid_phys_mins);
        o_phys2idx(cmax, cma, a->grid_step_size, a->grid_phys_mins);
        o_idx2phys(cmi, cminpos, a->grid_step_size, a->grid_phys_mins);
        for (int i = 0; i < 3; ++i)
            cgp[i] = cma[i] - cmi[i] + 1;
        int tot = interior_scan(cgp, cminpos, g_norm, g_f0, nf, a->grid_step_size,
                                a->eps);
        if (tot > 0) {
            if (tot > cap)
                return -1;
            float mpp = a->mass / (float)tot;      /* :1692 */
            int n = 0;
            for (int zi = 0; zi < cgp[2]; zi++)
                for (int yi = 0; yi < cgp[1]; yi++) {
                    int lo = g_border[2 * (zi * cgp[1] + yi)];
                    int hi = g_border[2 * (zi * cgp[1] + yi) + 1];
                    for (int xi = lo; xi <= hi; xi++) {
                        idx3[3 * n + 0] = cmi[0] + xi;
                        idx3[3 * n + 1] = cmi[1] + yi;
                        idx3[3 * n + 2] = cmi[2] + zi;
                        mass[n] = (double)mpp;
                        n++;
                    }
                }
            *status = O_CELL_SCAN;
            return n;
        }
        if (cap < 8)
            return -1;
        int i24[24];
        float m8[8];
        o_distribute_cic(&b->particles[3 * cell], a->mass, i24, m8,
                         a->grid_step_size, a->grid_phys_mins, a->eps);
        for (int i = 0; i < 8; ++i) {
            idx3[3 * i + 0] = i24[3 * i + 0];
            idx3[3 * i + 1] = i24[3 * i + 1];
            idx3[3 * i + 2] = i24[3 * i + 2];
            mass[i] = (double)m8[i];
        }
        *status = O_CELL_CIC;
        return 8;
    }
}

/* deposit one point (src/dense.cpp:276-305): in-block -> density,
 * else -> off-block list.  Returns 1 if off-block.  A point inside the
 * block bounds whose local index falls outside the block grid (possible on a
 * shared face, where BlockGridParams gives the index to the upper block; the
 * reference would index out of its row there) is treated as off-block. */
static int in_block_index(const o_block_t *b, const o_args_t *a, const int *bmin_idx,
                          const int *bnum_idx, const int *gidx)
{
    float pos[3];
    o_idx2phys(gidx, pos, a->grid_step_size, a->grid_phys_mins);
    int lo[3] = {gidx[0] - bmin_idx[0], gidx[1] - bmin_idx[1], gidx[2] - bmin_idx[2]};
    int ongrid = lo[0] >= 0 && lo[1] >= 0 && lo[0] < bnum_idx[0] && lo[1] < bnum_idx[1] &&
                 (a->project || (lo[2] >= 0 && lo[2] < bnum_idx[2]));
    if (ongrid && pos[0] >= b->bounds_min[0] && pos[0] <= b->bounds_max[0] &&
        pos[1] >= b->bounds_min[1] && pos[1] <= b->bounds_max[1] &&
        pos[2] >= b->bounds_min[2] && pos[2] <= b->bounds_max[2])
        return o_index(lo, bnum_idx, a->project, a->proj_plane);
    return -1;
}

static int deposit(const o_block_t *b, const o_args_t *a, const int *bmin_idx,
                   const int *bnum_idx, float *density, const int *gidx,
                   double m, float div, double *tot_mass)
{
    int idx = in_block_index(b, a, bmin_idx, bnum_idx, gidx);
    if (idx >= 0) {
        density[idx] = (float)((double)density[idx] + m / (double)div);
        if (tot_mass)
            *tot_mass += m;
        return 0;
    }
    return 1;
}

static int *g_pidx;
static double *g_pmass;
static int g_pcap;

long o_iterate_cells(const o_block_t *b, const o_args_t *a,
                     const int *block_min_idx, const int *block_num_idx,
                     float *density, int *cell_status, int *cell_npts,
                     int *off_idx3, double *off_mass, int *off_cell,
                     long off_cap, double *tot_mass)
{
    float div = a->project ? a->grid_step_size[0] * a->grid_step_size[1]
                           : a->grid_step_size[0] * a->grid_step_size[1] *
                                 a->grid_step_size[2];
    long noff = 0;
    ensure_faces(64);
    if (!g_pcap) {
        g_pcap = 4096;
        g_pidx = (int *)malloc(sizeof(int) * 3 * (size_t)g_pcap);
        g_pmass = (double *)malloc(sizeof(double) * (size_t)g_pcap);
    }
    for (int cell = 0; cell < b->num_orig_particles; cell++) {
        int st, n;
        while ((n = o_cell_points(b, a, cell, g_pidx, g_pmass, g_pcap, &st)) < 0) {
            g_pcap *= 4;
            g_pidx = (int *)realloc(g_pidx, sizeof(int) * 3 * (size_t)g_pcap);
            g_pmass = (double *)realloc(g_pmass, sizeof(double) * (size_t)g_pcap);
        }
        if (cell_status)
            cell_status[cell] = st;
        if (cell_npts)
            cell_npts[cell] = n;
        for (int i = 0; i < n; i++) {
            if (deposit(b, a, block_min_idx, block_num_idx, density, &g_pidx[3 * i],
                        g_pmass[i], div, tot_mass)) {
                if (noff < off_cap) {
                    off_idx3[3 * noff + 0] = g_pidx[3 * i + 0];
                    off_idx3[3 * noff + 1] = g_pidx[3 * i + 1];
                    off_idx3[3 * noff + 2] = g_pidx[3 * i + 2];
                    off_mass[noff] = g_pmass[i];
                    off_cell[noff] = cell;
                }
                noff++;
            }
        }
    }
    return noff;
}

long o_iterate_cells_fp64(const o_block_t *b, const o_args_t *a,
                          const int *block_min_idx, const int *block_num_idx,
                          double *density64)
{
    const double div = a->project
                           ? (double)a->grid_step_size[0] * (double)a->grid_step_size[1]
                           : (double)a->grid_step_size[0] * (double)a->grid_step_size[1] *
                                 (double)a->grid_step_size[2];
    long noff = 0;
    ensure_faces(64);
    if (!g_pcap) {
        g_pcap = 4096;
        g_pidx = (int *)malloc(sizeof(int) * 3 * (size_t)g_pcap);
        g_pmass = (double *)malloc(sizeof(double) * (size_t)g_pcap);
    }
    for (int cell = 0; cell < b->num_orig_particles; cell++) {
        int st, n;
        while ((n = o_cell_points(b, a, cell, g_pidx, g_pmass, g_pcap, &st)) < 0) {
            g_pcap *= 4;
            g_pidx = (int *)realloc(g_pidx, sizeof(int) * 3 * (size_t)g_pcap);
            g_pmass = (double *)realloc(g_pmass, sizeof(double) * (size_t)g_pcap);
        }
        /* the reference's mass / (float)N (:1692) taken in double; CIC
         * corners keep the reference's float weights */
        const double mscan = (double)a->mass / (double)n;
        for (int i = 0; i < n; i++) {
            int idx = in_block_index(b, a, block_min_idx, block_num_idx, &g_pidx[3 * i]);
            if (idx < 0) {
                noff++;
                continue;
            }
            const double m = st == O_CELL_SCAN ? mscan : g_pmass[i];
            density64[idx] += m / div;
        }
    }
    return noff;
}

long o_iterate_cells_cic(const o_block_t *b, const o_args_t *a,
                         const int *block_min_idx, const int *block_num_idx,
                         float *density, int *off_idx3, double *off_mass,
                         int *off_cell, long off_cap, double *tot_mass)
{
    float div = a->project ? a->grid_step_size[0] * a->grid_step_size[1]
                           : a->grid_step_size[0] * a->grid_step_size[1] *
                                 a->grid_step_size[2];
    long noff = 0;
    for (int cell = 0; cell < b->num_orig_particles; cell++) {
        int i24[24];
        float m8[8];
        o_distribute_cic(&b->particles[3 * cell], a->mass, i24, m8,
                         a->grid_step_size, a->grid_phys_mins, a->eps);
        for (int i = 0; i < 8; i++) {
            if (deposit(b, a, block_min_idx, block_num_idx, density, &i24[3 * i],
                        (double)m8[i], div, tot_mass)) {
                if (noff < off_cap) {
                    off_idx3[3 * noff + 0] = i24[3 * i + 0];
                    off_idx3[3 * noff + 1] = i24[3 * i + 1];
                    off_idx3[3 * noff + 2] = i24[3 * i + 2];
                    off_mass[noff] = (double)m8[i];
                    off_cell[noff] = cell;
                }
                noff++;
            }
        }
    }
    return noff;
}

void o_recvd_pts(const int *idx3, const double *mass, long n,
                 const int *block_min_idx, const int *block_num_idx,
                 const o_args_t *a, float *density)
{
    float div = a->project ? a->grid_step_size[0] * a->grid_step_size[1]
                           : a->grid_step_size[0] * a->grid_step_size[1] *
                                 a->grid_step_size[2];
    for (long j = 0; j < n; j++) {
        int l[3] = {idx3[3 * j] - block_min_idx[0], idx3[3 * j + 1] - block_min_idx[1],
                    idx3[3 * j + 2] - block_min_idx[2]};
        int idx = o_index(l, block_num_idx, a->project, a->proj_plane);
        density[idx] = (float)((double)density[idx] + mass[j] / (double)div);
    }
}

/* ------------------------------------------------------------------------ */
/* exact Voronoi owner (FP64 oracle)                                         */
/* ------------------------------------------------------------------------ */
static double d2_site(const float *g, const float *p)
{
    double dx = (double)g[0] - (double)p[0];
    double dy = (double)g[1] - (double)p[1];
    double dz = (double)g[2] - (double)p[2];
    return dx * dx + dy * dy + dz * dz;
}

void o_exact_owner(const float *particles, int np, const int *gnum,
                   const float *step, const float *gmin, int *owner)
{
    const int off[3] = {0, 0, 0};
    o_exact_owner_off(particles, np, off, gnum, step, gmin, owner);
}

void o_exact_owner_off(const float *particles, int np, const int *goff, const int *gnum,
                       const float *step, const float *gmin, int *owner)
{
    /* bucket hash over the particle bounding box, ~2 particles per bucket */
    double lo[3], hi[3];
    for (int i = 0; i < 3; ++i) {
        lo[i] = 1e300;
        hi[i] = -1e300;
    }
    for (int p = 0; p < np; ++p)
        for (int i = 0; i < 3; ++i) {
            double x = particles[3 * p + i];
            if (x < lo[i]) lo[i] = x;
            if (x > hi[i]) hi[i] = x;
        }
    double ext = 0.0;
    for (int i = 0; i < 3; ++i)
        if (hi[i] - lo[i] > ext) ext = hi[i] - lo[i];
    if (ext <= 0.0) ext = 1.0;
    int nb = (int)cbrt((double)np / 2.0);
    if (nb < 1) nb = 1;
    double h = ext / nb * (1.0 + 1e-9);
    int dims[3];
    for (int i = 0; i < 3; ++i) {
        dims[i] = (int)((hi[i] - lo[i]) / h) + 1;
    }
    long nbk = (long)dims[0] * dims[1] * dims[2];
    int *start = (int *)calloc((size_t)nbk + 1, sizeof(int));
    int *items = (int *)malloc(sizeof(int) * (size_t)np);
    int *bk = (int *)malloc(sizeof(int) * (size_t)np);
    for (int p = 0; p < np; ++p) {
        int c[3];
        for (int i = 0; i < 3; ++i) {
            c[i] = (int)(((double)particles[3 * p + i] - lo[i]) / h);
            if (c[i] >= dims[i]) c[i] = dims[i] - 1;
        }
        bk[p] = (c[2] * dims[1] + c[1]) * dims[0] + c[0];
        start[bk[p] + 1]++;
    }
    for (long i = 0; i < nbk; ++i)
        start[i + 1] += start[i];
    int *fill = (int *)malloc(sizeof(int) * (size_t)nbk);
    memcpy(fill, start, sizeof(int) * (size_t)nbk);
    for (int p = 0; p < np; ++p)       /* ascending id within each bucket */
        items[fill[bk[p]]++] = p;
    free(fill);
    free(bk);

    for (int z = 0; z < gnum[2]; ++z)
        for (int y = 0; y < gnum[1]; ++y)
            for (int x = 0; x < gnum[0]; ++x) {
                int gi[3] = {x + goff[0], y + goff[1], z + goff[2]};
                float g[3];
                o_idx2phys(gi, g, step, gmin);
                int c[3];
                for (int i = 0; i < 3; ++i) {
                    double f = ((double)g[i] - lo[i]) / h;
                    int ci = (int)floor(f);
                    if (ci < 0) ci = 0;
                    if (ci >= dims[i]) ci = dims[i] - 1;
                    c[i] = ci;
                }
                double best = 1e300;
                int bid = -1;
                for (int r = 0;; ++r) {
                    /* shell r around c: buckets with Chebyshev distance r */
                    int any = 0;
                    for (int dz = -r; dz <= r; ++dz)
                        for (int dy = -r; dy <= r; ++dy)
                            for (int dx = -r; dx <= r; ++dx) {
                                if (abs(dx) != r && abs(dy) != r && abs(dz) != r)
                                    continue;
                                int q[3] = {c[0] + dx, c[1] + dy, c[2] + dz};
                                if (q[0] < 0 || q[1] < 0 || q[2] < 0 || q[0] >= dims[0] ||
                                    q[1] >= dims[1] || q[2] >= dims[2])
                                    continue;
                                any = 1;
                                long bb = ((long)q[2] * dims[1] + q[1]) * dims[0] + q[0];
                                for (int s = start[bb]; s < start[bb + 1]; ++s) {
                                    int p = items[s];
                                    double d = d2_site(g, &particles[3 * p]);
                                    if (d < best || (d == best && p < bid)) {
                                        best = d;
                                        bid = p;
                                    }
                                }
                            }
                    /* every bucket beyond shell r lies at least r*h away
                     * (g may sit outside the hash box: then distances are
                     * only larger); stop once that exceeds the best */
                    double lb = (double)r * h;
                    if (bid >= 0 && lb * lb > best * (1.0 + 1e-9) + 1e-30)
                        break;
                    if (!any && r > dims[0] + dims[1] + dims[2])
                        break;
                }
                owner[((long)z * gnum[1] + y) * gnum[0] + x] = bid;
            }
    free(start);
    free(items);
}

/* ------------------------------------------------------------------------ */
/* exact-Voronoi density (FP64 oracle)                                       */
/* ------------------------------------------------------------------------ */
static void cc_f64(const o_tet_t *t, const float *P, double *o)
{
    const float *a = &P[3 * t->verts[0]], *b = &P[3 * t->verts[1]];
    const float *c = &P[3 * t->verts[2]], *d = &P[3 * t->verts[3]];
    double t0 = (double)a[0] - d[0], t1 = (double)a[1] - d[1], t2 = (double)a[2] - d[2];
    double u0 = (double)b[0] - d[0], u1 = (double)b[1] - d[1], u2 = (double)b[2] - d[2];
    double v0 = (double)c[0] - d[0], v1 = (double)c[1] - d[1], v2 = (double)c[2] - d[2];
    double nt = t0 * t0 + t1 * t1 + t2 * t2;
    double nu = u0 * u0 + u1 * u1 + u2 * u2;
    double nv = v0 * v0 + v1 * v1 + v2 * v2;
    double uv0 = u1 * v2 - u2 * v1, uv1 = u2 * v0 - u0 * v2, uv2 = u0 * v1 - u1 * v0;
    double vt0 = v1 * t2 - v2 * t1, vt1 = v2 * t0 - v0 * t2, vt2 = v0 * t1 - v1 * t0;
    double tu0 = t1 * u2 - t2 * u1, tu1 = t2 * u0 - t0 * u2, tu2 = t0 * u1 - t1 * u0;
    double den = 2.0 * (t0 * uv0 + t1 * uv1 + t2 * uv2);
    o[0] = (double)d[0] + (nt * uv0 + nu * vt0 + nv * tu0) / den;
    o[1] = (double)d[1] + (nt * uv1 + nu * vt1 + nv * tu1) / den;
    o[2] = (double)d[2] + (nt * uv2 + nu * vt2 + nv * tu2) / den;
}

void o_exact_dense(const o_block_t *b, const o_args_t *a, const int *bmin_idx,
                   const int *bnum, int *owner, int *status, int *npts, double *density)
{
    const int nc = b->num_orig_particles;
    float lo[3], hi[3];
    for (int i = 0; i < 3; ++i) {
        float e = (a->data_maxs[i] - a->data_mins[i]) * 2.0f * FLT_EPSILON;
        lo[i] = a->data_mins[i] - e;
        hi[i] = a->data_maxs[i] + e;
    }
    /* 1. which cells deposit */
    for (int c = 0; c < nc; ++c) {
        int vt = b->vert_to_tet[c];
        if (vt < 0 || !star_bfs(c, b->tets, vt, &g_q, &g_vis, &g_vv, NULL, NULL)) {
            status[c] = O_CELL_INCOMPLETE;
            continue;
        }
        /* flat (coplanar, cocircular) tets have no circumcenter: skipped */
        double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
        int defined = 0;
        for (int k = 0; k < g_vis.n; ++k) {
            double o[3];
            cc_f64(&b->tets[g_vis.p[k]], b->particles, o);
            if (isnan(o[0]) || isnan(o[1]) || isnan(o[2]))
                continue;
            defined++;
            for (int i = 0; i < 3; ++i) {
                if (o[i] < mn[i]) mn[i] = o[i];
                if (o[i] > mx[i]) mx[i] = o[i];
            }
        }
        if (!defined)
            status[c] = O_CELL_DEGENERATE;
        else if (mn[0] < lo[0] || mn[1] < lo[1] || mn[2] < lo[2] || mx[0] > hi[0] ||
                 mx[1] > hi[1] || mx[2] > hi[2])
            status[c] = O_CELL_OUTSIDE;
        else
            status[c] = O_CELL_SCAN;
    }
    /* 2. nearest site of every block grid point */
    const long ng = (long)bnum[0] * bnum[1] * bnum[2];
    o_exact_owner_off(b->particles, b->num_particles, bmin_idx, bnum, a->grid_step_size,
                      a->grid_phys_mins, owner);
    for (int c = 0; c < nc; ++c)
        npts[c] = 0;
    for (long g = 0; g < ng; ++g) {
        int x = (int)(g % bnum[0]), y = (int)((g / bnum[0]) % bnum[1]),
            z = (int)(g / ((long)bnum[0] * bnum[1]));
        int gi[3] = {bmin_idx[0] + x, bmin_idx[1] + y, bmin_idx[2] + z};
        float pos[3];
        o_idx2phys(gi, pos, a->grid_step_size, a->grid_phys_mins);
        if (!(pos[0] >= b->bounds_min[0] && pos[0] <= b->bounds_max[0] &&
              pos[1] >= b->bounds_min[1] && pos[1] <= b->bounds_max[1] &&
              pos[2] >= b->bounds_min[2] && pos[2] <= b->bounds_max[2])) {
            owner[g] = -1;
            continue;
        }
        int o = owner[g];
        if (o >= 0 && o < nc && status[o] == O_CELL_SCAN)
            npts[o]++;
    }
    /* 3. density: contributions in ascending cell order */
    const double div = (double)(a->grid_step_size[0] * a->grid_step_size[1] *
                                a->grid_step_size[2]);
    for (long g = 0; g < ng; ++g)
        density[g] = 0.0;
    /* owner contributions: each point has at most one */
    double *own_part = (double *)calloc((size_t)ng, sizeof(double));
    int *own_cell = (int *)malloc(sizeof(int) * (size_t)ng);
    for (long g = 0; g < ng; ++g) {
        int o = owner[g];
        own_cell[g] = -1;
        if (o >= 0 && o < nc && status[o] == O_CELL_SCAN && npts[o] > 0) {
            own_part[g] = (double)a->mass / (double)npts[o];
            own_cell[g] = o;
        }
    }
    /* CIC contributions of deposited cells without points, in cell order,
     * merged with the owner contribution at its cell position */
    for (int c = 0; c < nc; ++c) {
        if (status[c] != O_CELL_SCAN || npts[c] > 0)
            continue;
        status[c] = O_CELL_CIC;
        npts[c] = 8;
        int i24[24];
        float m8[8];
        o_distribute_cic(&b->particles[3 * c], a->mass, i24, m8, a->grid_step_size,
                         a->grid_phys_mins, a->eps);
        for (int k = 0; k < 8; ++k) {
            float pos[3];
            o_idx2phys(&i24[3 * k], pos, a->grid_step_size, a->grid_phys_mins);
            if (!(pos[0] >= b->bounds_min[0] && pos[0] <= b->bounds_max[0] &&
                  pos[1] >= b->bounds_min[1] && pos[1] <= b->bounds_max[1] &&
                  pos[2] >= b->bounds_min[2] && pos[2] <= b->bounds_max[2]))
                continue;
            int l[3] = {i24[3 * k] - bmin_idx[0], i24[3 * k + 1] - bmin_idx[1],
                        i24[3 * k + 2] - bmin_idx[2]};
            if (l[0] < 0 || l[1] < 0 || l[2] < 0 || l[0] >= bnum[0] || l[1] >= bnum[1] ||
                l[2] >= bnum[2])
                continue;
            long g = ((long)l[2] * bnum[1] + l[1]) * bnum[0] + l[0];
            if (own_cell[g] >= 0 && own_cell[g] < c) {
                density[g] += own_part[g];
                own_cell[g] = -2;     /* folded */
            }
            density[g] += (double)m8[k];
        }
    }
    for (long g = 0; g < ng; ++g) {
        if (own_cell[g] >= 0)
            density[g] += own_part[g];
        density[g] /= div;
    }
    free(own_part);
    free(own_cell);
}

/* ------------------------------------------------------------------------ */
/* ghost selection: incomplete_cells(), src/tess.cpp:492-628                */
/* ------------------------------------------------------------------------ */

/* dot(u, p, x) = u . (p - x), accumulated from 0, src/tet.cpp:21-29 */
static float o_dot3(const float *u, const float *p, const float *x)
{
    float res = 0.0f;
    for (int i = 0; i < 3; ++i)
        res += u[i] * (p[i] - x[i]);
    return res;
}

/* side_of_plane(), src/tet.cpp:70-119: does a point of the box lie on the
 * other side of the facet opposite vertex j than vertex j itself?  A flat
 * facet (vertex j on its plane) counts as yes. */
static int o_side_of_plane(const float *bmin, const float *bmax, const o_tet_t *tet,
                           const float *particles, int j)
{
    int idx[4] = {0, 1, 2, 3};
    int tmp = idx[j];
    idx[j] = idx[3];
    idx[3] = tmp;
    const float *x = &particles[3 * tet->verts[idx[0]]];
    const float *y = &particles[3 * tet->verts[idx[1]]];
    const float *z = &particles[3 * tet->verts[idx[2]]];
    const float *p = &particles[3 * tet->verts[idx[3]]];
    float u[3], v[3], n[3];
    for (int i = 0; i < 3; ++i) {
        u[i] = y[i] - x[i];
        v[i] = z[i] - x[i];
    }
    o_cross(n, u, v);
    float res = o_dot3(n, p, x);
    int sign = 0;
    if (res > 0)
        sign = 1;
    else if (res < 0)
        sign = -1;
    if (sign == 0)
        return 1;
    res = 0.0f;
    for (int i = 0; i < 3; ++i) {
        if ((sign > 0) != (n[i] < 0))
            res += n[i] * (bmin[i] - x[i]);
        else
            res += n[i] * (bmax[i] - x[i]);
    }
    return (sign > 0) != (res > 0);
}

/* Distance from a point to a box, 0 inside: per axis the larger of
 * (min - p) and (p - max) when positive, squared and summed in float, then
 * the square root.  This is DIY's distance(Bounds, Point), used at
 * src/tess.cpp:568; DIY is not part of the reference tree (an unpinned
 * GitHub checkout, .build.yml:11-17), so the float rounding of this helper
 * is restated from its published source, not pinned by a reference build. */
static float o_box_distance(const float *bmin, const float *bmax, const float *p)
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

void o_ghost_sends(const o_block_t *b, int nnb, const float *boxes, int first,
                   unsigned long long *sends)
{
    const int norig = b->num_orig_particles;
    for (int p = 0; p < norig; ++p)
        sends[p] = 0ull;
    for (int t = 0; t < b->num_tets; t++) {
        const o_tet_t *tet = &b->tets[t];
        int j;
        for (j = 0; j < 4; ++j)
            if (tet->verts[j] < norig)
                break;
        if (j == 4)     /* no local particle in the tet */
            continue;
        float center[3];
        o_circumcenter(center, tet, b->particles);
        const float rad = o_distance(center, &b->particles[3 * tet->verts[0]]);
        /* convex hull facets with a local particle (:516-547) */
        for (j = 0; j < 4; ++j) {
            if (tet->tets[j] != -1)
                continue;
            int k;
            for (k = 0; k < 4; ++k) {
                if (k == j)
                    continue;
                if (tet->verts[k] < norig)
                    break;
            }
            if (k == 4)
                continue;
            for (int i = first; i < nnb; ++i) {
                if (!o_side_of_plane(&boxes[6 * i], &boxes[6 * i + 3], tet, b->particles, j))
                    continue;
                for (int q = 0; q < 4; q++) {
                    if (q == j)
                        continue;
                    const int pp = tet->verts[q];
                    if (pp < norig)
                        sends[pp] |= 1ull << i;
                }
            }
        }
        /* a circumsphere too deep inside the block cannot stick out (:550-557) */
        for (j = 0; j < 3; ++j) {
            if (center[j] - b->bounds_min[j] <= rad)
                break;
            if (b->bounds_max[j] - center[j] <= rad)
                break;
        }
        if (j == 3)
            continue;
        /* neighbour boxes within the circumradius (:560-578) */
        for (int i = first; i < nnb; ++i) {
            if (o_box_distance(&boxes[6 * i], &boxes[6 * i + 3], center) <= rad) {
                for (int q = 0; q < 4; q++) {
                    const int pp = tet->verts[q];
                    if (pp < norig)
                        sends[pp] |= 1ull << i;
                }
            }
        }
    }
}
