// grid_io.cpp -- WriteGrid() and ProjectGrid() for one process holding every
// block (src/dense.cpp:751-1023): the 2-D projection's sum of z-stacked block
// planes into the z = 0 block of the same x-y column, and the raw float32
// output file with each block's grid at its place in the global array.
// Host code: the output is a file, and the projection sum runs once per
// write over planes the caller already holds.
#include <cstdio>
#include <cstring>
#include <vector>

#include <fcntl.h>
#include <unistd.h>

#include "../../include/tess2_amd.h"

namespace {

struct BlockRef {
    int gid;
    float *density;
    int mn[3], num[3];
};

// ProjectGrid(), src/dense.cpp:881-1023, all blocks in one process: blocks in
// ascending gid order; every block whose grid does not start at z index 0
// adds its plane, float += float point by point, into the block whose grid
// has the same x-y start and starts at z index 0 (the root).  The root's
// plane and the adding block's plane have the same x-y extent (same x-y
// start, block grids tile the domain).
int project_blocks(std::vector<BlockRef> &b)
{
    for (const BlockRef &src : b) {
        if (src.mn[2] == 0) continue;
        const BlockRef *root = nullptr;
        for (const BlockRef &r : b)
            if (r.mn[0] == src.mn[0] && r.mn[1] == src.mn[1] && r.mn[2] == 0) root = &r;
        if (!root) return TESS2_E_ARG;
        if (root->num[0] != src.num[0] || root->num[1] != src.num[1]) return TESS2_E_ARG;
        const size_t n = (size_t)src.num[0] * (size_t)src.num[1];
        for (size_t i = 0; i < n; ++i) root->density[i] += src.density[i];
    }
    return TESS2_OK;
}

int collect(int nblocks, const int *gids, float **density, const int *block_min_idx,
            const int *block_num_idx, std::vector<BlockRef> &b)
{
    if (nblocks <= 0 || !gids || !density || !block_min_idx || !block_num_idx)
        return TESS2_E_ARG;
    b.resize((size_t)nblocks);
    for (int i = 0; i < nblocks; ++i) {
        BlockRef &r = b[(size_t)i];
        r.gid = gids[i];
        r.density = density[i];
        for (int k = 0; k < 3; ++k) {
            r.mn[k] = block_min_idx[3 * i + k];
            r.num[k] = block_num_idx[3 * i + k];
            if (r.mn[k] < 0 || r.num[k] < 0) return TESS2_E_ARG;
        }
        if (!r.density && (size_t)r.num[0] * r.num[1] * r.num[2] > 0) return TESS2_E_ARG;
    }
    // ascending gid: the order a single process's master holds its blocks
    for (size_t i = 1; i < b.size(); ++i)
        for (size_t j = i; j > 0 && b[j - 1].gid > b[j].gid; --j) std::swap(b[j - 1], b[j]);
    return TESS2_OK;
}

bool write_at(int fd, const float *p, size_t n, size_t off_elems)
{
    const char *c = reinterpret_cast<const char *>(p);
    size_t left = n * sizeof(float);
    off_t off = (off_t)(off_elems * sizeof(float));
    while (left) {
        const ssize_t w = pwrite(fd, c, left, off);
        if (w <= 0) return false;
        c += w;
        off += w;
        left -= (size_t)w;
    }
    return true;
}

}  // namespace

extern "C" int tess2_project_grid(int nblocks, const int *gids, float **density,
                                  const int *block_min_idx, const int *block_num_idx)
{
    std::vector<BlockRef> b;
    int r = collect(nblocks, gids, density, block_min_idx, block_num_idx, b);
    if (r) return r;
    return project_blocks(b);
}

extern "C" int tess2_write_blocks(const char *path, int nblocks, const int *gids,
                                  float **density, const int *block_min_idx,
                                  const int *block_num_idx, const int *glo_num_idx, int project)
{
    if (!path || !glo_num_idx) return TESS2_E_ARG;
    std::vector<BlockRef> b;
    int r = collect(nblocks, gids, density, block_min_idx, block_num_idx, b);
    if (r) return r;
    const size_t gx = (size_t)glo_num_idx[0], gy = (size_t)glo_num_idx[1];
    for (const BlockRef &k : b)
        for (int a = 0; a < 3; ++a)
            if ((size_t)k.mn[a] + (size_t)k.num[a] > (size_t)glo_num_idx[a]) return TESS2_E_ARG;
    if (project && (r = project_blocks(b))) return r;
    // start from an empty file (MPI_File_set_size(fd, 0), src/dense.cpp:785)
    const int fd = open(path, O_WRONLY | O_CREAT | O_TRUNC, 0644);
    if (fd < 0) return TESS2_E_IO;
    bool ok = true;
    for (const BlockRef &k : b) {
        const size_t nx = (size_t)k.num[0], ny = (size_t)k.num[1], nz = (size_t)k.num[2];
        if (project) {
            // [y][x] plane; blocks not at z index 0 write nothing (:822-823)
            if (k.mn[2] != 0) continue;
            for (size_t y = 0; ok && y < ny; ++y)
                ok = write_at(fd, k.density + y * nx, nx, (k.mn[1] + y) * gx + k.mn[0]);
        } else {
            for (size_t z = 0; ok && z < nz; ++z)
                for (size_t y = 0; ok && y < ny; ++y)
                    ok = write_at(fd, k.density + (z * ny + y) * nx, nx,
                                  ((k.mn[2] + z) * gy + (k.mn[1] + y)) * gx + k.mn[0]);
        }
        if (!ok) break;
    }
    if (close(fd) != 0) ok = false;
    return ok ? TESS2_OK : TESS2_E_IO;
}
