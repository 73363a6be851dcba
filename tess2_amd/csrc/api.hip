// api.hip -- the C ABI (include/tess2_amd.h): context, staging of blocks in
// HBM, the run sequence on one HIP stream, and the reference's host-side
// grid helpers.
#include "../../include/tess2_amd.h"
#include "geom.hpp"
#include "kernels.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

using namespace t2;

// ---------------------------------------------------------------------------
// Device memory of the allocation-free entry point (tess2_volumes_noalloc):
// one file-scope array, which every GPU the library runs on holds from the
// code object's load until the process ends.  Host code reaches it only
// through the address hipGetSymbolAddress returns for the current device
// (scratch_ptr); the array's name on the host is not a device address.
// ---------------------------------------------------------------------------
static constexpr size_t NOALLOC_SCRATCH_BYTES = 64ull << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

static std::mutex g_scratch_mu;
static std::vector<unsigned char *> g_scratch_ptr;   // per device; nullptr until looked up

// once per device, with that device current (tess2_create)
static hipError_t scratch_lookup(int device)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if ((int)g_scratch_ptr.size() <= device) g_scratch_ptr.resize((size_t)device + 1, nullptr);
    if (g_scratch_ptr[(size_t)device]) return hipSuccess;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e == hipSuccess) g_scratch_ptr[(size_t)device] = static_cast<unsigned char *>(p);
    return e;
}

static unsigned char *scratch_ptr(int device)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return device >= 0 && device < (int)g_scratch_ptr.size() ? g_scratch_ptr[(size_t)device]
                                                             : nullptr;
}

namespace {

struct DevBlock {
    int gid = 0, norig = 0, np = 0, nt = 0;
    float bmin[3] = {0, 0, 0}, bmax[3] = {0, 0, 0};
    float4 *pts = nullptr;
    int4 *tets = nullptr;
    int *v2t = nullptr;
    float4 *cc = nullptr;
    int *order = nullptr;                    // cells in Morton order
    std::vector<int> tet_new;                // caller's tet id -> device tet id
    float *vol = nullptr;
    bool vol_valid = false;                  // the last run wrote vol
    int *status = nullptr, *npts = nullptr;
    // grid
    GridP gp{};
    size_t ngrid = 0, grid_cap = 0;
    float *density = nullptr;
    unsigned *count = nullptr;
    uint2 *first = nullptr;
    uint2 *second = nullptr;
    int *head = nullptr;
    int4 *ovf = nullptr;
    unsigned ovf_cap = 0;
    uint2 *seg = nullptr;                 // fold segments, 3 * ovf_cap
    int4 *fold[3] = {nullptr, nullptr, nullptr};
    unsigned fold_cap[3] = {0, 0, 0};
    int *slow = nullptr;
    unsigned slow_cap = 0;
    // exact-Voronoi assignment
    int *owner = nullptr;                    // [grid] nearest site
    int *status_ex = nullptr;                // [ncells] deposit decision
    unsigned *exflags = nullptr;             // [3 * ncells bytes] its per-vertex flags
    HashP hash{};
    unsigned *hstart = nullptr, *hfill = nullptr, *hsums = nullptr;
    int *hbk = nullptr;
    float4 *hsorted = nullptr;
    int *hkey = nullptr;                     // [np] home key of each particle (several blocks)
    std::vector<int> ghost_gid, ghost_cell;  // home of each ghost, when given
    size_t owner_cap = 0;
    unsigned *defq = nullptr;                // FP32 filter failures awaiting the FP64 search
    unsigned defq_cap = 0;
    BlockStats *st = nullptr;                // this block's record in the context's array
    unsigned *ctrs = nullptr;                // st->ctrs
    BlockStats *hm = nullptr;                // its pinned host copy, refreshed by every run
    int block_max_idx[3] = {0, 0, 0};
};

struct RunArgs {
    int alg = 0, num_given = 0, project = 0, assignment = 0, precision = 0, want_vol = 1;
    float given_mins[3] = {0, 0, 0}, given_maxs[3] = {0, 0, 0};
    float mass = 1.0f, eps = 1e-4f;
    int gnum[3] = {0, 0, 0};
};

}  // namespace

struct tess2_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    std::vector<DevBlock> blocks;
    hipEvent_t ev_run0 = nullptr, ev_run1 = nullptr, ev_k0 = nullptr, ev_k1 = nullptr;
    RunArgs last;
    bool have_run = false;
    float data_mins[3] = {0, 0, 0}, data_maxs[3] = {0, 0, 0};
    float gmin[3] = {0, 0, 0}, gmax[3] = {0, 0, 0}, step[3] = {0, 0, 0};
    double ktime_sum = 0.0;
    int ktime_n = 0;
    float last_run_ms = 0.0f;
    // multi-block: descriptors of all staged blocks on the device
    int ncu = 256;
    BlockIn *d_ins = nullptr;
    GridP *d_gps = nullptr;
    CellOut *d_outs = nullptr;
    Router *d_router = nullptr;
    int *d_rmap = nullptr;              // the router's interval maps and table
    int2 *d_rcand = nullptr;
    size_t rmap_cap = 0, rcand_cap = 0;
    Chunk *d_chunks = nullptr;
    int nchunks = 0;
    Chunk *d_items = nullptr;            // two consecutive cells of one block each
    int nitems = 0;
    XcdQueues ixr{};                     // the item list's XCD ranges
    unsigned *d_queue = nullptr;
    XcdQueues xq{};
    std::vector<BlockIn> h_ins;
    std::vector<GridP> h_gps;
    std::vector<CellOut> h_outs;
    // batched schedule (tess2_autotune): chunk size per block size class and
    // persistent waves per CU
    int chunk_small = 16, chunk_medium = 32, chunk_large = 64, waves_per_cu = 24;
    // single-block cell kernel: -1 = chosen per run (cell_kernel_for), 0 =
    // one cell per wave, 1 / 2 = two cells per wave in the geometry phase
    // (launch_cells_mc), 2 with the multi-word verdict masks
    int cell_kernel = -1;
    // graph replay of the steady-state sequence (tess2_set_option)
    bool graph_on = false;
    hipGraphExec_t gexec = nullptr;
    RunArgs gargs;
    unsigned long long gen = 1;      // bumped whenever a run buffer is reallocated
    unsigned long long ggen = 0;     // generation the graph was captured at
    bool in_capture = false;
    BlockStats *d_stats = nullptr, *h_stats = nullptr;   // per-block counters, device / pinned
    double *d_gpart = nullptr;                           // partial grid sums, 1024 per block
    // exact assignment over several blocks: every particle's home key
    // (rank << 23 | cell), and the per-rank npts / deposit-status tables
    bool homes_ok = false;
    std::string homes_err;
    int **d_npts_tab = nullptr;
    int **d_stex_tab = nullptr;
    bool kevents = false;            // the last run recorded the kernel events
    RunArgs desc_args;
    unsigned long long desc_gen = 0;
};

static bool same_args(const RunArgs &x, const RunArgs &y)
{
    return memcmp(&x, &y, sizeof(RunArgs)) == 0;
}

// The single-block cell kernel of a run: the option when set, else two cells
// per wave, with the multi-word verdict masks when cell boxes are large
// (16 or more grid points per cell: 2.1M cells on 512^3, 20.2 vs 24.7 ms;
// at 8 points per cell the masks cost 2 %).
static int cell_kernel_for(const tess2_ctx *c, int ncells, unsigned long long ngrid)
{
    // the multi-block choices; one block: one cell per wave, or two with
    // the multi-word masks
    if (c->cell_kernel == 3) return 0;
    if (c->cell_kernel == 4) return 2;
    if (c->cell_kernel >= 0) return c->cell_kernel;
    return ngrid >= 16ull * (unsigned long long)std::max(ncells, 1) ? 2 : 1;
}

static void drop_graph(tess2_ctx *c)
{
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    c->gexec = nullptr;
    c->ggen = 0;
}

static void free_multi(tess2_ctx *c)
{
    void *ps[] = {c->d_ins, c->d_gps, c->d_outs, c->d_router, c->d_chunks, c->d_queue,
                  c->d_stats, c->d_gpart, c->d_npts_tab, c->d_stex_tab, c->d_rmap, c->d_rcand,
                  c->d_items};
    c->d_items = nullptr;
    c->nitems = 0;
    c->d_rmap = nullptr;
    c->d_rcand = nullptr;
    c->rmap_cap = c->rcand_cap = 0;
    c->d_npts_tab = nullptr;
    c->d_stex_tab = nullptr;
    c->homes_ok = false;
    for (void *p : ps)
        if (p) (void)hipFree(p);
    if (c->h_stats) (void)hipHostFree(c->h_stats);
    c->d_stats = nullptr;
    c->h_stats = nullptr;
    c->d_gpart = nullptr;
    c->d_ins = nullptr;
    c->d_gps = nullptr;
    c->d_outs = nullptr;
    c->d_router = nullptr;
    c->d_chunks = nullptr;
    c->d_queue = nullptr;
    c->nchunks = 0;
}

static int fail(tess2_ctx *c, int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    return code;
}

#define HIPCHK(c, x)                                                                   \
    do {                                                                               \
        hipError_t e_ = (x);                                                           \
        if (e_ != hipSuccess)                                                          \
            return fail((c), TESS2_E_HIP, "%s failed: %s", #x, hipGetErrorString(e_)); \
    } while (0)

template <class T>
static int dalloc(tess2_ctx *c, T **p, size_t n)
{
    if (*p) {
        (void)hipFree(*p);
        *p = nullptr;
    }
    if (n == 0) n = 1;
    if (c) c->gen++;   // any captured graph refers to the old buffers
    hipError_t e = hipMalloc((void **)p, n * sizeof(T));
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(c, TESS2_E_NOMEM, "hipMalloc(%zu bytes): %s", n * sizeof(T),
                    hipGetErrorString(e));
    }
    return 0;
}

static void free_block(DevBlock &b)
{
    void *ps[] = {b.order,  b.pts,    b.tets,   b.v2t,   b.cc,   b.vol,  b.status,    b.npts,
                  b.density, b.count, b.first, b.second, b.head, b.ovf,  b.slow,
                  b.owner, b.status_ex, b.exflags, b.hstart, b.hfill,
                  b.hsums,  b.hbk,    b.hsorted, b.hkey, b.defq, b.seg, b.fold[0], b.fold[1],
                  b.fold[2]};
    for (void *p : ps)
        if (p) (void)hipFree(p);
    b = DevBlock();
}

// ---------------------------------------------------------------------------
// host helpers: GridStepParams / BlockGridParams (same arithmetic as the
// reference, see geom.hpp for the float conventions)
// ---------------------------------------------------------------------------
extern "C" void tess2_grid_step_params(int num_given_bounds, const float *given_mins,
                                       const float *given_maxs, const float *data_mins,
                                       const float *data_maxs, float *gmin, float *gmax,
                                       float *step, const int *gnum)
{
    // src/dense.cpp:1721-1766; the padding is divided by the double 2.0
    float mds = data_maxs[0] - data_mins[0];
    if (data_maxs[1] - data_mins[1] > mds) mds = data_maxs[1] - data_mins[1];
    if (data_maxs[2] - data_mins[2] > mds) mds = data_maxs[2] - data_mins[2];
    for (int i = 0; i < 3; ++i) {
        const float pad = mds - (data_maxs[i] - data_mins[i]);
        gmin[i] = (float)((double)data_mins[i] - (double)pad / 2.0);
        gmax[i] = (float)((double)data_maxs[i] + (double)pad / 2.0);
    }
    for (int i = 0; i < 3; ++i)
        if (num_given_bounds >= i + 1) {
            gmin[i] = given_mins[i];
            gmax[i] = given_maxs[i];
        }
    for (int i = 0; i < 3; ++i) step[i] = (gmax[i] - gmin[i]) / (float)(gnum[i] - 1);
}

extern "C" void tess2_block_grid_params(const float *bmin, const float *bmax, int *mn,
                                        int *mx, int *num, const float *gmin,
                                        const float *step, float eps, const float *data_mins,
                                        const float *data_maxs, const int *gnum)
{
    // src/dense.cpp:586-647
    for (int i = 0; i < 3; ++i) {
        mn[i] = phys2idx1(bmin[i], step[i], gmin[i]);
        if (idx2phys1(mn[i], step[i], gmin[i]) < bmin[i]) mn[i]++;
        mx[i] = phys2idx1(bmax[i], step[i], gmin[i]);
        if (idx2phys1(mx[i], step[i], gmin[i]) + step[i] <= bmax[i]) mx[i]++;
    }
    for (int i = 0; i < 3; ++i)
        if (std::fabs(data_mins[i] + (float)mx[i] * step[i] - bmax[i]) < eps &&
            std::fabs(bmax[i] - data_maxs[i]) > step[i])
            mx[i]--;
    for (int i = 0; i < 3; ++i)
        if (std::fabs(bmin[i] - data_mins[i]) < step[i]) mn[i] = 0;
    for (int i = 0; i < 3; ++i)
        if (std::fabs(bmax[i] - data_maxs[i]) < step[i]) mx[i] = gnum[i] - 1;
    for (int i = 0; i < 3; ++i) num[i] = mx[i] - mn[i] + 1;
}

// ---------------------------------------------------------------------------
extern "C" int tess2_create(tess2_ctx **out, int device)
{
    if (!out) return TESS2_E_ARG;
    *out = nullptr;
    tess2_ctx *c = new tess2_ctx();
    c->device = device;
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) {
        delete c;
        return TESS2_E_HIP;
    }
    // default priority, non-blocking w.r.t. the null stream
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev_run0) != hipSuccess || hipEventCreate(&c->ev_run1) != hipSuccess ||
        hipEventCreate(&c->ev_k0) != hipSuccess || hipEventCreate(&c->ev_k1) != hipSuccess) {
        delete c;
        return TESS2_E_HIP;
    }
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess &&
        ncu > 0)
        c->ncu = ncu;
    // the allocation-free entry point's scratch address on this device; a
    // failed lookup only disables that entry point (it reports the error)
    const hipError_t se = scratch_lookup(device);
    if (se != hipSuccess)
        fail(c, TESS2_E_HIP, "scratch lookup on device %d: %s", device, hipGetErrorString(se));
    *out = c;
    return 0;
}

extern "C" int tess2_destroy(tess2_ctx *c)
{
    if (!c) return 0;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    drop_graph(c);
    for (auto &b : c->blocks) free_block(b);
    free_multi(c);
    if (c->ev_run0) (void)hipEventDestroy(c->ev_run0);
    if (c->ev_run1) (void)hipEventDestroy(c->ev_run1);
    if (c->ev_k0) (void)hipEventDestroy(c->ev_k0);
    if (c->ev_k1) (void)hipEventDestroy(c->ev_k1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return 0;
}

extern "C" const char *tess2_last_error(const tess2_ctx *c)
{
    return c ? c->err.c_str() : "null context";
}

// ---------------------------------------------------------------------------
// staging
// ---------------------------------------------------------------------------
static inline uint32_t spread3(uint32_t x)
{
    x &= 0x3ffu;
    x = (x | (x << 16)) & 0x030000FFu;
    x = (x | (x << 8)) & 0x0300F00Fu;
    x = (x | (x << 4)) & 0x030C30C3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}

// Morton ranks of all particles, the original cells in Morton order, and the
// tets renumbered by the smallest Morton rank among their vertices.
static void morton_layout(const tess2_block_t &hb, std::vector<int> &rank,
                          std::vector<int> &order, std::vector<int> &tnew,
                          std::vector<tess2_tet_t> &tl, std::vector<int> &v2t)
{
    const int np = hb.num_particles, nt = hb.num_tets;
    float lo[3] = {0, 0, 0}, hi[3] = {1, 1, 1};
    for (int p = 0; p < np; ++p)
        for (int i = 0; i < 3; ++i) {
            const float x = hb.particles[3 * p + i];
            if (p == 0 || x < lo[i]) lo[i] = x;
            if (p == 0 || x > hi[i]) hi[i] = x;
        }
    std::vector<uint64_t> key(np);
    for (int p = 0; p < np; ++p) {
        uint32_t q[3];
        for (int i = 0; i < 3; ++i) {
            const double ext = (double)hi[i] - lo[i];
            double f = ext > 0 ? ((double)hb.particles[3 * p + i] - lo[i]) / ext : 0.0;
            q[i] = (uint32_t)std::min(1023.0, std::max(0.0, f * 1023.0));
        }
        const uint32_t code = spread3(q[0]) | (spread3(q[1]) << 1) | (spread3(q[2]) << 2);
        key[p] = ((uint64_t)code << 32) | (uint32_t)p;
    }
    std::sort(key.begin(), key.end());
    rank.assign(np, 0);
    order.clear();
    order.reserve(hb.num_orig_particles);
    for (int k = 0; k < np; ++k) {
        const int p = (int)(key[k] & 0xffffffffu);
        rank[p] = k;
        if (p < hb.num_orig_particles) order.push_back(p);
    }
    std::vector<uint64_t> tk(nt);
    for (int t = 0; t < nt; ++t) {
        int m = rank[hb.tets[t].verts[0]];
        for (int i = 1; i < 4; ++i) m = std::min(m, rank[hb.tets[t].verts[i]]);
        tk[t] = ((uint64_t)(uint32_t)m << 32) | (uint32_t)t;
    }
    std::sort(tk.begin(), tk.end());
    tnew.assign(nt, 0);
    for (int k = 0; k < nt; ++k) tnew[(int)(tk[k] & 0xffffffffu)] = k;
    tl.resize(nt);
    for (int k = 0; k < nt; ++k) {
        const tess2_tet_t &q = hb.tets[(int)(tk[k] & 0xffffffffu)];
        tess2_tet_t &o = tl[k];
        for (int i = 0; i < 4; ++i) {
            o.verts[i] = q.verts[i];
            o.tets[i] = q.tets[i] < 0 ? q.tets[i] : tnew[q.tets[i]];
        }
    }
    v2t.resize(np);
    for (int p = 0; p < np; ++p)
        v2t[p] = hb.vert_to_tet[p] < 0 ? hb.vert_to_tet[p] : tnew[hb.vert_to_tet[p]];
}

static int validate(tess2_ctx *c, const tess2_block_t &b)
{
    if (b.num_orig_particles < 0 || b.num_particles < b.num_orig_particles || b.num_tets < 0)
        return fail(c, TESS2_E_ARG, "block %d: bad counts", b.gid);
    if ((b.num_particles && !b.particles) || (b.num_tets && !b.tets) ||
        (b.num_particles && !b.vert_to_tet))
        return fail(c, TESS2_E_ARG, "block %d: null array", b.gid);
    const int np = b.num_particles, nt = b.num_tets;
    for (int t = 0; t < nt; ++t) {
        const tess2_tet_t &q = b.tets[t];
        for (int i = 0; i < 4; ++i) {
            if (q.verts[i] < 0 || q.verts[i] >= np)
                return fail(c, TESS2_E_ARG, "block %d: tet %d vertex %d out of range", b.gid, t,
                            q.verts[i]);
            if (q.tets[i] < -1 || q.tets[i] >= nt)
                return fail(c, TESS2_E_ARG, "block %d: tet %d neighbour %d out of range", b.gid,
                            t, q.tets[i]);
        }
    }
    for (int p = 0; p < np; ++p)
        if (b.vert_to_tet[p] < -1 || b.vert_to_tet[p] >= nt)
            return fail(c, TESS2_E_ARG, "block %d: vert_to_tet[%d] out of range", b.gid, p);
    return 0;
}

// Work list of the batched launch: each block's cells (in its Morton order)
// cut into chunks whose size depends on the block's size class.
static int build_chunks(tess2_ctx *c)
{
    // blocks in staging (gid) order, each block's cells in Morton order, so
    // that the 8 XCD ranges below are compact in space
    std::vector<Chunk> ch;
    long long total = 0;
    for (const DevBlock &d : c->blocks) total += d.norig;
    // Chunks interleave the cells of a tile of G * cs cells over G chunks
    // (chunk i: cells i, i + G, ...), G = the waves of one XCD, so that the
    // cells an XCD works on at any time stay within about G consecutive ones
    // (its L2 working set) rather than G whole chunks (64 Plummer blocks:
    // 17.9 -> 17.7 ms; 8 uniform blocks 7.22 -> 7.07 ms)
    const int G = std::max(1, c->ncu * c->waves_per_cu / 8);
    for (int b = 0; b < (int)c->blocks.size(); ++b) {
        const int n = c->blocks[b].norig;
        const int cs = n < 20000 ? c->chunk_small : (n < 100000 ? c->chunk_medium : c->chunk_large);
        for (int t0 = 0; t0 < n; t0 += G * cs) {
            const int m = std::min(G * cs, n - t0);
            const int g = (m + cs - 1) / cs;
            for (int i = 0; i < g; ++i) ch.push_back(Chunk{b, t0 + i, (m - i + g - 1) / g, g});
        }
    }
    c->nchunks = (int)ch.size();
    // split into 8 ranges of about equal cell counts
    long long acc = 0;
    int x = 1;
    c->xq.start[0] = 0;
    for (int i = 0; i < (int)ch.size() && x < 8; ++i) {
        acc += ch[i].count;
        while (x < 8 && acc * 8 >= total * x) c->xq.start[x++] = i + 1;
    }
    while (x <= 8) c->xq.start[x++] = (int)ch.size();
    int r = dalloc(c, &c->d_chunks, std::max<size_t>(ch.size(), 1));
    if (r) return r;
    if (!ch.empty())
        HIPCHK(c, hipMemcpy(c->d_chunks, ch.data(), sizeof(Chunk) * ch.size(),
                            hipMemcpyHostToDevice));
    // the item list of the queue-free launches: two cells at a time
    std::vector<Chunk> it;
    for (int b = 0; b < (int)c->blocks.size(); ++b)
        for (int s0 = 0; s0 < c->blocks[b].norig; s0 += 2)
            it.push_back(Chunk{b, s0, std::min(2, c->blocks[b].norig - s0), 1});
    c->nitems = (int)it.size();
    if ((r = dalloc(c, &c->d_items, std::max<size_t>(it.size(), 1)))) return r;
    if (!it.empty())
        HIPCHK(c, hipMemcpy(c->d_items, it.data(), sizeof(Chunk) * it.size(),
                            hipMemcpyHostToDevice));
    return 0;
}

// Home keys for the exact assignment over several blocks: particle p of the
// block at rank k is (k << 23 | p) when it is an original, else its home
// block's rank and cell.  Without a complete, consistent home for every ghost
// the exact assignment over several blocks is refused at run time (the
// reference scan does not need it).
static int build_home_keys(tess2_ctx *c)
{
    const int nb = (int)c->blocks.size();
    c->homes_ok = false;
    std::vector<int> gids(nb);
    for (int k = 0; k < nb; ++k) gids[k] = c->blocks[k].gid;
    for (int k = 0; k < nb; ++k) {
        const DevBlock &d = c->blocks[k];
        if (d.np > d.norig && d.ghost_gid.empty()) {
            c->homes_err = "block " + std::to_string(d.gid) + " gives no ghost homes";
            return 0;
        }
    }
    std::vector<int> key;
    for (int k = 0; k < nb; ++k) {
        DevBlock &d = c->blocks[k];
        key.resize(d.np);
        for (int p = 0; p < d.norig; ++p) key[p] = (k << 23) | p;
        for (int p = d.norig; p < d.np; ++p) {
            const int g = d.ghost_gid[p - d.norig], cell = d.ghost_cell[p - d.norig];
            const auto it = std::lower_bound(gids.begin(), gids.end(), g);
            if (it == gids.end() || *it != g) {
                c->homes_err = "a ghost of block " + std::to_string(d.gid) + " names block " +
                               std::to_string(g) + ", which is not staged";
                return 0;
            }
            const int r = (int)(it - gids.begin());
            if (cell < 0 || cell >= c->blocks[r].norig) {
                c->homes_err = "a ghost of block " + std::to_string(d.gid) +
                               " names a cell outside block " + std::to_string(g);
                return 0;
            }
            key[p] = (r << 23) | cell;
        }
        int rr = 0;
        if ((rr = dalloc(c, &d.hkey, (size_t)d.np))) return rr;
        if (d.np)
            HIPCHK(c, hipMemcpy(d.hkey, key.data(), sizeof(int) * (size_t)d.np,
                                hipMemcpyHostToDevice));
    }
    std::vector<int *> tn(nb), ts(nb);
    for (int k = 0; k < nb; ++k) {
        tn[k] = c->blocks[k].npts;
        ts[k] = c->blocks[k].status_ex;
    }
    int rr = 0;
    if ((rr = dalloc(c, &c->d_npts_tab, (size_t)nb)) || (rr = dalloc(c, &c->d_stex_tab, (size_t)nb)))
        return rr;
    HIPCHK(c, hipMemcpy(c->d_npts_tab, tn.data(), sizeof(int *) * nb, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_stex_tab, ts.data(), sizeof(int *) * nb, hipMemcpyHostToDevice));
    c->homes_ok = true;
    return 0;
}

extern "C" int tess2_stage(tess2_ctx *c, const tess2_block_t *blocks, int nblocks)
{
    if (!c || (nblocks > 0 && !blocks) || nblocks < 0) return fail(c, TESS2_E_ARG, "bad args");
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int i = 0; i < nblocks; ++i) {
        int r = validate(c, blocks[i]);
        if (r) return r;
    }
    // deposit keys: cell << 1 | z bit for own cells; 1 << 31 | rank << 24 |
    // cell << 1 | z bit for points received from another block
    for (int i = 0; i < nblocks; ++i)
        if (blocks[i].num_orig_particles >= (1 << 30))
            return fail(c, TESS2_E_ARG, "at most 2^30 - 1 cells per block");
    if (nblocks > 1) {
        if (nblocks > 127) return fail(c, TESS2_E_ARG, "at most 127 blocks per context");
        for (int i = 0; i < nblocks; ++i) {
            if (blocks[i].num_orig_particles >= (1 << 23))
                return fail(c, TESS2_E_ARG, "multi-block runs need < 2^23 cells per block");
            if (i > 0 && blocks[i].gid <= blocks[i - 1].gid)
                return fail(c, TESS2_E_ARG, "blocks must be given in ascending gid order");
        }
    }
    drop_graph(c);
    c->gen++;
    for (auto &b : c->blocks) free_block(b);
    free_multi(c);
    c->blocks.assign(nblocks, DevBlock());
    c->have_run = false;
    std::vector<float4> tmp;
    for (int k = 0; k < nblocks; ++k) {
        const tess2_block_t &hb = blocks[k];
        DevBlock &d = c->blocks[k];
        d.gid = hb.gid;
        d.norig = hb.num_orig_particles;
        d.np = hb.num_particles;
        d.nt = hb.num_tets;
        if (hb.ghost_gid && hb.ghost_cell && d.np > d.norig) {
            d.ghost_gid.assign(hb.ghost_gid, hb.ghost_gid + (d.np - d.norig));
            d.ghost_cell.assign(hb.ghost_cell, hb.ghost_cell + (d.np - d.norig));
        }
        for (int i = 0; i < 3; ++i) {
            d.bmin[i] = hb.bounds_min[i];
            d.bmax[i] = hb.bounds_max[i];
        }
        int r = 0;
        if ((r = dalloc(c, &d.pts, (size_t)d.np)) || (r = dalloc(c, &d.tets, 2 * (size_t)d.nt)) ||
            (r = dalloc(c, &d.v2t, (size_t)d.np)) || (r = dalloc(c, &d.cc, (size_t)d.nt)) ||
            (r = dalloc(c, &d.vol, (size_t)d.norig)) || (r = dalloc(c, &d.status, (size_t)d.norig)) ||
            (r = dalloc(c, &d.npts, (size_t)d.norig)))
            return r;
        tmp.resize(d.np);
        for (int p = 0; p < d.np; ++p)
            tmp[p] = make_float4(hb.particles[3 * p], hb.particles[3 * p + 1],
                                 hb.particles[3 * p + 2], 0.0f);
        HIPCHK(c, hipMemcpy(d.pts, tmp.data(), sizeof(float4) * d.np, hipMemcpyHostToDevice));
        // HBM layout: cells visited in Morton order of their sites; tets
        // renumbered by the Morton rank of their first-ranked vertex so that
        // the stars of cells processed together are close in memory.  The
        // results do not depend on either order (the star traversal follows
        // neighbour slots, not tet ids).
        std::vector<int> rank, order, tnew;
        std::vector<tess2_tet_t> tl;
        std::vector<int> v2t;
        morton_layout(hb, rank, order, tnew, tl, v2t);
        d.tet_new.swap(tnew);
        if ((r = dalloc(c, &d.order, (size_t)d.norig))) return r;
        if (d.norig)
            HIPCHK(c, hipMemcpy(d.order, order.data(), sizeof(int) * (size_t)d.norig,
                                hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(d.tets, tl.data(), sizeof(tess2_tet_t) * (size_t)d.nt,
                            hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(d.v2t, v2t.data(), sizeof(int) * (size_t)d.np,
                            hipMemcpyHostToDevice));
        d.slow_cap = (unsigned)std::max(1024, d.norig / 32);
        if ((r = dalloc(c, &d.slow, d.slow_cap))) return r;
        // bucket hash over the particle bounding box, about two particles
        // per bucket; at most 2^24 - 1024 buckets, so that the bucket scan
        // needs at most 16384 blocks of 1024 (k_scan_sums' limit)
        float lo[3] = {0, 0, 0}, hi[3] = {1, 1, 1};
        for (int p = 0; p < d.np; ++p)
            for (int i = 0; i < 3; ++i) {
                const float x = hb.particles[3 * p + i];
                if (p == 0 || x < lo[i]) lo[i] = x;
                if (p == 0 || x > hi[i]) hi[i] = x;
            }
        double vol = 1.0;
        for (int i = 0; i < 3; ++i) vol *= std::max((double)hi[i] - lo[i], 1e-12);
        double cell = std::cbrt(vol * 2.0 / std::max(d.np, 1));
        for (;;) {
            long long nb = 1;
            for (int i = 0; i < 3; ++i)
                nb *= (long long)((hi[i] - lo[i]) / cell) + 1;
            if (nb <= (1ll << 24) - 1024) break;
            cell *= 1.25;
        }
        d.hash.cell = (float)cell;
        d.hash.inv = (float)(1.0 / cell);
        for (int i = 0; i < 3; ++i) {
            d.hash.lo[i] = lo[i];
            d.hash.dims[i] = (int)((hi[i] - lo[i]) / cell) + 1;
        }
        const size_t nbk = (size_t)d.hash.dims[0] * d.hash.dims[1] * d.hash.dims[2];
        if ((r = dalloc(c, &d.hstart, nbk + 1)) || (r = dalloc(c, &d.hfill, nbk + 1)) ||
            (r = dalloc(c, &d.hsums, (nbk + 1 + 1023) / 1024)) ||
            (r = dalloc(c, &d.hbk, (size_t)d.np)) || (r = dalloc(c, &d.hsorted, (size_t)d.np)) ||
            (r = dalloc(c, &d.status_ex, (size_t)d.norig)) ||
            (r = dalloc(c, &d.exflags, ((size_t)d.norig * 3 + 3) / 4)))
            return r;
    }
    if (nblocks > 0) {
        int r = 0;
        if ((r = dalloc(c, &c->d_ins, (size_t)nblocks)) || (r = dalloc(c, &c->d_gps, (size_t)nblocks)) ||
            (r = dalloc(c, &c->d_outs, (size_t)nblocks)) || (r = dalloc(c, &c->d_router, 1)) ||
            (r = dalloc(c, &c->d_queue, 16)) || (r = dalloc(c, &c->d_stats, (size_t)nblocks)) ||
            (r = dalloc(c, &c->d_gpart, (size_t)nblocks * 1024)))
            return r;
        // counters and statistics of all blocks: one device array, one pinned
        // host copy (so that the copy can sit in a captured graph)
        if (hipHostMalloc((void **)&c->h_stats, sizeof(BlockStats) * nblocks,
                          hipHostMallocDefault) != hipSuccess)
            return fail(c, TESS2_E_NOMEM, "pinned host buffer");
        memset(c->h_stats, 0, sizeof(BlockStats) * nblocks);
        for (int k = 0; k < nblocks; ++k) {
            DevBlock &d = c->blocks[k];
            d.st = c->d_stats + k;
            d.ctrs = d.st->ctrs;
            d.hm = c->h_stats + k;
        }
        r = build_chunks(c);
        if (r) return r;
        if (nblocks > 1 && (r = build_home_keys(c))) return r;
    }
    return 0;
}

// shared-point list and the fold's segment storage and class lists for
// `cap` list entries (at most 2 * cap contributions in shared points)
static int alloc_lists(tess2_ctx *c, DevBlock &d, unsigned cap)
{
    int r = 0;
    // segments hold a point's list entries plus its first and second
    // contributions: at most 3 * cap pairs; the list doubles as the large
    // fold's scratch of the same size
    const size_t c3 = 3 * (size_t)cap;
    const unsigned c1 = (unsigned)(c3 / (FOLD_SMALL_MAX + 1) + 1),
                   c2 = (unsigned)(c3 / (FOLD_MID_MAX + 1) + 1);
    if ((r = dalloc(c, &d.ovf, (c3 + 1) / 2)) || (r = dalloc(c, &d.seg, c3)) ||
        (r = dalloc(c, &d.fold[0], cap)) || (r = dalloc(c, &d.fold[1], c1)) ||
        (r = dalloc(c, &d.fold[2], c2)))
        return r;
    d.ovf_cap = cap;
    d.fold_cap[0] = cap;
    d.fold_cap[1] = c1;
    d.fold_cap[2] = c2;
    return 0;
}

// grid + list buffers for one run
static int ensure_grid(tess2_ctx *c, DevBlock &d, size_t n)
{
    d.ngrid = n;
    if (n > d.grid_cap) {
        int r = 0;
        if ((r = dalloc(c, &d.density, n)) || (r = dalloc(c, &d.count, n)) ||
            (r = dalloc(c, &d.first, n)) || (r = dalloc(c, &d.second, n)) ||
            (r = dalloc(c, &d.head, n)))
            return r;
        d.grid_cap = n;
    }
    if (!d.ovf) {
        int r = alloc_lists(c, d, (unsigned)std::min<size_t>(std::max<size_t>(1u << 16, n / 8),
                                                             1u << 30));
        if (r) return r;
    }
    return 0;
}

static CellOut make_out(DevBlock &d, bool want_vol)
{
    CellOut o;
    o.vol = want_vol ? d.vol : nullptr;
    o.status = d.status;
    o.npts = d.npts;
    o.density = d.density;
    o.count = d.count;
    o.first = d.first;
    o.second = d.second;
    o.head = d.head;
    o.ovf = d.ovf;
    o.ovf_cap = d.ovf_cap;
    o.seg = d.seg;
    for (int k = 0; k < 3; ++k) {
        o.fold[k] = d.fold[k];
        o.fold_cap[k] = d.fold_cap[k];
    }
    o.st = d.st;
    o.ngrid = d.ngrid;
    o.ncells = d.norig;
    o.ctrs = d.ctrs;
    o.slow = d.slow;
    o.slow_cap = d.slow_cap;
    return o;
}

static BlockIn make_in(DevBlock &d)
{
    BlockIn in;
    in.pts = d.pts;
    in.tets = d.tets;
    in.v2t = d.v2t;
    in.cc = d.cc;
    in.order = d.order;
    in.ncells = d.norig;
    in.ntets = d.nt;
    return in;
}

// The router's tables (Router::rmap / rcand) from the blocks' 3-D grid
// ranges: a point that a cell deposits outside its own block goes to the
// first block (staging order) other than its own whose range holds it
// (src/dense.cpp:298-305).  The range ends cut each axis into intervals; a
// box of intervals lies in a block's range entirely or not at all, so its
// first two holders answer the question for every sender.  At most 127
// blocks: at most 253 intervals per axis.
static int build_route_table(tess2_ctx *c, Router &hr)
{
    const int nb = (int)c->blocks.size();
    std::vector<int> cuts[3];
    for (int a = 0; a < 3; ++a) {
        for (int k = 0; k < nb; ++k) {
            cuts[a].push_back(c->h_gps[k].bmin_idx[a]);
            cuts[a].push_back(c->h_gps[k].bmin_idx[a] + c->h_gps[k].bnum_idx[a]);
        }
        std::sort(cuts[a].begin(), cuts[a].end());
        cuts[a].erase(std::unique(cuts[a].begin(), cuts[a].end()), cuts[a].end());
        hr.rbase[a] = cuts[a].front();
        hr.rlen[a] = cuts[a].back() - cuts[a].front();
        hr.rn[a] = (int)cuts[a].size() - 1;
    }
    hr.roff[0] = 0;
    hr.roff[1] = hr.rlen[0];
    hr.roff[2] = hr.rlen[0] + hr.rlen[1];
    std::vector<int> map((size_t)hr.rlen[0] + hr.rlen[1] + hr.rlen[2]);
    for (int a = 0; a < 3; ++a)
        for (int i = 0; i < hr.rn[a]; ++i)
            for (int g = cuts[a][i]; g < cuts[a][i + 1]; ++g)
                map[(size_t)hr.roff[a] + (g - hr.rbase[a])] = i;
    const size_t n1 = (size_t)hr.rn[1], n2 = (size_t)hr.rn[2];
    std::vector<int2> cand((size_t)hr.rn[0] * n1 * n2, make_int2(-1, -1));
    for (int k = 0; k < nb; ++k) {
        int lo[3], hi[3];
        for (int a = 0; a < 3; ++a) {
            const int b0 = c->h_gps[k].bmin_idx[a], b1 = b0 + c->h_gps[k].bnum_idx[a];
            lo[a] = (int)(std::lower_bound(cuts[a].begin(), cuts[a].end(), b0) - cuts[a].begin());
            hi[a] = (int)(std::lower_bound(cuts[a].begin(), cuts[a].end(), b1) - cuts[a].begin());
        }
        for (int i0 = lo[0]; i0 < hi[0]; ++i0)
            for (int i1 = lo[1]; i1 < hi[1]; ++i1)
                for (int i2 = lo[2]; i2 < hi[2]; ++i2) {
                    int2 &e = cand[((size_t)i0 * n1 + i1) * n2 + i2];
                    if (e.x < 0) e.x = k;
                    else if (e.y < 0) e.y = k;
                }
    }
    int r = 0;
    if (map.size() > c->rmap_cap) {
        if ((r = dalloc(c, &c->d_rmap, map.size()))) return r;
        c->rmap_cap = map.size();
    }
    if (cand.size() > c->rcand_cap) {
        if ((r = dalloc(c, &c->d_rcand, cand.size()))) return r;
        c->rcand_cap = cand.size();
    }
    // on the library stream, after any earlier run that still reads the
    // tables; the host vectors must outlive the copies
    HIPCHK(c, hipMemcpyAsync(c->d_rmap, map.data(), sizeof(int) * map.size(),
                             hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->d_rcand, cand.data(), sizeof(int2) * cand.size(),
                             hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    hr.rmap = c->d_rmap;
    hr.rcand = c->d_rcand;
    return 0;
}

// enqueue the whole path for every staged block
// host-side set-up of a run: grid parameters, buffers, device descriptors
static int prepare_run(tess2_ctx *c, const RunArgs &a)
{
    if (a.gnum[0] < 2 || a.gnum[1] < 2 || a.gnum[2] < 2)
        return fail(c, TESS2_E_ARG, "grid needs at least 2 points per axis");
    if (c->blocks.empty()) return fail(c, TESS2_E_ARG, "no staged blocks");
    const int nb = (int)c->blocks.size();
    const bool exact = a.alg == TESS2_DENSE_TESS && a.assignment == TESS2_ASSIGN_EXACT_VORONOI;
    if (exact && nb > 1 && !c->homes_ok)
        return fail(c, TESS2_E_ARG, "the exact assignment over several blocks: %s",
                    c->homes_err.c_str());
    // DataBounds(): min/max over the block bounds (src/dense.cpp:1238-1267)
    for (int k = 0; k < nb; ++k)
        for (int i = 0; i < 3; ++i) {
            const DevBlock &b = c->blocks[k];
            if (k == 0 || b.bmin[i] < c->data_mins[i]) c->data_mins[i] = b.bmin[i];
            if (k == 0 || b.bmax[i] > c->data_maxs[i]) c->data_maxs[i] = b.bmax[i];
        }
    tess2_grid_step_params(a.num_given, a.given_mins, a.given_maxs, c->data_mins, c->data_maxs,
                           c->gmin, c->gmax, c->step, a.gnum);
    const float div = a.project ? c->step[0] * c->step[1] : c->step[0] * c->step[1] * c->step[2];
    hipStream_t s = c->stream;

    // per-block grids and buffers
    c->h_ins.resize(nb);
    c->h_gps.resize(nb);
    c->h_outs.resize(nb);
    for (int k = 0; k < nb; ++k) {
        DevBlock &d = c->blocks[k];
        GridP &gp = d.gp;
        int mx[3];
        tess2_block_grid_params(d.bmin, d.bmax, gp.bmin_idx, mx, gp.bnum_idx, c->gmin, c->step,
                                a.eps, c->data_mins, c->data_maxs, a.gnum);
        for (int i = 0; i < 3; ++i) {
            d.block_max_idx[i] = mx[i];
            gp.gmin[i] = c->gmin[i];
            gp.step[i] = c->step[i];
            gp.dmin[i] = c->data_mins[i];
            gp.dmax[i] = c->data_maxs[i];
            gp.bmin[i] = d.bmin[i];
            gp.bmax[i] = d.bmax[i];
        }
        if (gp.bnum_idx[0] <= 0 || gp.bnum_idx[1] <= 0 || gp.bnum_idx[2] <= 0)
            return fail(c, TESS2_E_ARG, "block %d has an empty grid", d.gid);
        gp.eps = a.eps;
        gp.mass = a.mass;
        gp.div = div;
        gp.project = a.project;
        const size_t n = (size_t)gp.bnum_idx[0] * gp.bnum_idx[1] *
                         (a.project ? 1 : (size_t)gp.bnum_idx[2]);
        int r = ensure_grid(c, d, n);
        if (r) return r;
        if (exact && d.owner_cap < n) {
            r = dalloc(c, &d.owner, n);
            if (r) return r;
            d.owner_cap = n;
            // a quarter of the grid (beyond it the search runs inline); the
            // queue holds 32-bit grid indices
            if (d.defq) (void)hipFree(d.defq);
            d.defq = nullptr;
            d.defq_cap = 0;
            if (n < (1ull << 32)) {
                const size_t cap = n / 4 + 1024;
                r = dalloc(c, &d.defq, cap);
                if (r) return r;
                d.defq_cap = (unsigned)cap;
            }
        }
        c->h_ins[k] = make_in(d);
        c->h_gps[k] = gp;
        c->h_outs[k] = make_out(d, a.want_vol != 0);
    }
    if (c->desc_gen == c->gen && same_args(a, c->desc_args)) return 0;   // unchanged
    if (nb > 1) {
        // XCD ranges of the item list by estimated cost: a cell, plus its
        // share of the block's grid points at 0.006 cells each (a least-
        // squares fit of the per-block item times of this launch on 64
        // Plummer blocks; ranges by cell count there: 24.1 ms, by cost 21.3)
        const double beta = 0.006;
        std::vector<double> cum(c->nitems + 1, 0.0);
        int it = 0;
        for (int k = 0; k < nb; ++k) {
            const DevBlock &d = c->blocks[k];
            const int ni = (d.norig + 1) / 2;
            const double per = ni ? (d.norig + beta * (double)d.ngrid) / ni : 0.0;
            for (int i = 0; i < ni; ++i, ++it) cum[it + 1] = cum[it] + per;
        }
        c->ixr.start[0] = 0;
        int x = 1;
        for (int i = 1; i <= c->nitems && x < 8; ++i)
            while (x < 8 && cum[i] * 8.0 >= cum[c->nitems] * x) c->ixr.start[x++] = i;
        while (x <= 8) c->ixr.start[x++] = c->nitems;
    }
    Router hr{};
    hr.nblocks = nb;
    hr.gps = c->d_gps;
    hr.outs = c->d_outs;
    if (nb > 1) {
        int r = build_route_table(c, hr);
        if (r) return r;
    }
    HIPCHK(c, hipMemcpyAsync(c->d_ins, c->h_ins.data(), sizeof(BlockIn) * nb,
                             hipMemcpyHostToDevice, s));
    HIPCHK(c, hipMemcpyAsync(c->d_gps, c->h_gps.data(), sizeof(GridP) * nb,
                             hipMemcpyHostToDevice, s));
    HIPCHK(c, hipMemcpyAsync(c->d_outs, c->h_outs.data(), sizeof(CellOut) * nb,
                             hipMemcpyHostToDevice, s));
    HIPCHK(c, hipMemcpyAsync(c->d_router, &hr, sizeof(Router), hipMemcpyHostToDevice, s));
    // the host copies must outlive the asynchronous transfers
    HIPCHK(c, hipStreamSynchronize(s));
    c->desc_gen = c->gen;
    c->desc_args = a;
    return 0;
}

// the kernel sequence of a run (captured into a graph in graph mode)
static int enqueue_kernels(tess2_ctx *c, const RunArgs &a)
{
    const int nb = (int)c->blocks.size();
    const bool exact = a.alg == TESS2_DENSE_TESS && a.assignment == TESS2_ASSIGN_EXACT_VORONOI;
    hipStream_t s = c->stream;
    const bool ev = !c->in_capture;

    // phase A: clear per-block outputs and counters, circumcenters
    launch_fill_u32(reinterpret_cast<unsigned *>(c->d_stats), sizeof(BlockStats) / 4 * nb, 0u, s);
    {
        int mt = 0;
        unsigned long long mg = 0;
        for (const DevBlock &d : c->blocks) {
            mt = std::max(mt, d.nt);
            mg = std::max<unsigned long long>(mg, d.ngrid);
        }
        launch_block_prep(c->d_ins, c->d_outs, nb, mt, mg, a.alg == TESS2_DENSE_TESS, s);
    }
    // phase B: the estimator
    if (exact) {
        // several blocks: particle ids are home keys, every block's cell
        // counts are complete before any density is written (all statuses,
        // then all counts, then the densities), and off-block CIC corners
        // are routed like the reference scan's points
        const bool multi = nb > 1;
        int *const *ntab = multi ? c->d_npts_tab : nullptr;
        const int *const *stab = multi ? c->d_stex_tab : nullptr;
        for (int k = 0; k < nb; ++k) {
            DevBlock &d = c->blocks[k];
            const BlockIn &in = c->h_ins[k];
            const GridP &gp = c->h_gps[k];
            if (a.want_vol) {
                // volume() as the reference computes it
                CellOut ov = make_out(d, true);
                const int ck = nb > 1 ? 0 : cell_kernel_for(c, d.norig, d.ngrid);
                if (ck == 0)
                    launch_cells_ref(in, gp, ov, 0, s);
                else if (launch_cells_mc(in, c->d_gps + k, ov, 0, ck, s))
                    return fail(c, TESS2_E_ARG, "unknown cell kernel %d", ck);
                launch_cells_ref_slow(in, gp, ov, 0, s);
            }
            launch_hash_build(d.pts, d.np, d.hash, d.hstart, d.hfill, d.hsums, d.hbk, d.hsorted,
                              multi ? d.hkey : nullptr, s);
            launch_exact_status(in, gp, d.status_ex, a.precision, d.ctrs,
                                reinterpret_cast<unsigned char *>(d.exflags), s);
            if (multi && d.norig > 0) launch_fill_u32((unsigned *)d.npts, (size_t)d.norig, 0u, s);
        }
        if (ev) HIPCHK(c, hipEventRecord(c->ev_k0, s));
        for (int k = 0; k < nb; ++k) {
            DevBlock &d = c->blocks[k];
            launch_owner(c->h_gps[k], d.hash, d.hstart, d.hsorted, d.owner, a.precision, d.ctrs,
                         d.status_ex, d.npts, d.norig, ntab, stab, d.defq, d.defq_cap, s);
        }
        if (ev) HIPCHK(c, hipEventRecord(c->ev_k1, s));
        // every density pass before any CIC pass: an owned point's density is
        // its first contribution, stored without atomics, and CIC corners
        // routed to another block must come after it.  A cell's count changes
        // in its CIC pass only when it owns no point, so no density pass
        // would read a rewritten count anyway.
        for (int part = 1; part <= 2; ++part)
            for (int k = 0; k < nb; ++k)
                launch_exact_tail(c->h_ins[k], c->h_gps[k], c->blocks[k].owner,
                                  c->blocks[k].status_ex, c->h_outs[k], a.precision, ntab, stab,
                                  multi ? c->d_router : nullptr, k, part, s);
    } else if (a.alg == TESS2_DENSE_CIC) {
        if (ev) HIPCHK(c, hipEventRecord(c->ev_k0, s));
        for (int k = 0; k < nb; ++k)
            launch_cic(c->h_ins[k], c->h_gps[k], c->h_outs[k], nb > 1 ? c->d_router : nullptr, k,
                       s);
        if (ev) HIPCHK(c, hipEventRecord(c->ev_k1, s));
    } else if (nb == 1) {
        if (ev) HIPCHK(c, hipEventRecord(c->ev_k0, s));
        const int ck = cell_kernel_for(c, c->blocks[0].norig, c->blocks[0].ngrid);
        if (ck == 0)
            launch_cells_ref(c->h_ins[0], c->h_gps[0], c->h_outs[0], 1, s);
        else if (launch_cells_mc(c->h_ins[0], c->d_gps, c->h_outs[0], 1, ck, s))
            return fail(c, TESS2_E_ARG, "unknown cell kernel %d", ck);
        if (ev) HIPCHK(c, hipEventRecord(c->ev_k1, s));
        launch_cells_ref_slow(c->h_ins[0], c->h_gps[0], c->h_outs[0], 1, s);
    } else {
        if (ev) HIPCHK(c, hipEventRecord(c->ev_k0, s));
        // one cell per wave, the star search's record loads issued before
        // the table work (option 0: after it, in a gather round per level),
        // unless a two-cell kernel is asked for: on 64 Plummer blocks the
        // two-cell kernel measured 24.1 vs 22.7 ms (its 48-tet capacity
        // defers 6,241 cells of the dense cores) and the early loads 23.6 vs
        // 24.0 ms
        if (c->cell_kernel == 4) {
            launch_cells_mc_items(c->d_ins, c->d_gps, c->d_outs, c->d_router, c->d_items,
                                  c->ixr, 1, s);
            launch_cells_ref_slow_batched(c->d_ins, c->d_router, 1, s);
        } else if (c->cell_kernel <= 0 || c->cell_kernel == 3) {
            launch_cells_ref_batched(c->d_ins, c->d_gps, c->d_outs, c->d_router, c->d_chunks,
                                     c->xq, c->d_queue, c->ncu * c->waves_per_cu, 1, s,
                                     c->cell_kernel != 0);
        } else {
            launch_cells_mc_batched(c->d_ins, c->d_gps, c->d_outs, c->d_router, c->d_chunks,
                                    c->xq, c->d_queue, c->ncu * c->waves_per_cu, 1, s);
            launch_cells_ref_slow_batched(c->d_ins, c->d_router, 1, s);
        }
        if (ev) HIPCHK(c, hipEventRecord(c->ev_k1, s));
    }
    // phase C: ordered fold of shared points, statistics, one copy of the
    // counters of all blocks
    unsigned mcap = 0, mfold[3] = {0, 0, 0};
    int mcells = 0;
    unsigned long long mgrid = 0, sovf = 0, smid = 0;
    for (const DevBlock &d : c->blocks) {
        mcap = std::max(mcap, d.ovf_cap);
        sovf += d.ovf_cap;
        smid += d.fold_cap[1];
        for (int k = 0; k < 3; ++k) mfold[k] = std::max(mfold[k], d.fold_cap[k]);
        mcells = std::max(mcells, d.norig);
        mgrid = std::max<unsigned long long>(mgrid, d.ngrid);
    }
    launch_fold(c->d_outs, c->d_gps, nb, mcap, mfold, mgrid, sovf, smid, s);
    launch_stats(c->d_outs, c->d_gpart, nb, mcells, mgrid, s);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(c->h_stats, c->d_stats, sizeof(BlockStats) * nb,
                             hipMemcpyDeviceToHost, s));
    return 0;
}

static int enqueue_run(tess2_ctx *c, const RunArgs &a)
{
    int r = prepare_run(c, a);
    if (r) return r;
    // the cell kernels write volume() only in a TESS run that asks for it
    for (DevBlock &d : c->blocks)
        d.vol_valid = a.alg == TESS2_DENSE_TESS && a.want_vol != 0;
    hipStream_t s = c->stream;
    if (c->graph_on) {
        if (!c->gexec || c->ggen != c->gen || !same_args(a, c->gargs)) {
            drop_graph(c);
            hipGraph_t g = nullptr;
            HIPCHK(c, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            c->in_capture = true;
            r = enqueue_kernels(c, a);
            c->in_capture = false;
            hipError_t e = hipStreamEndCapture(s, &g);
            if (r) {
                if (g) (void)hipGraphDestroy(g);
                return r;
            }
            if (e != hipSuccess) return fail(c, TESS2_E_HIP, "graph capture: %s", hipGetErrorString(e));
            e = hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) {
                c->gexec = nullptr;
                return fail(c, TESS2_E_HIP, "graph instantiate: %s", hipGetErrorString(e));
            }
            c->ggen = c->gen;
            c->gargs = a;
        }
        HIPCHK(c, hipEventRecord(c->ev_run0, s));
        HIPCHK(c, hipGraphLaunch(c->gexec, s));
        HIPCHK(c, hipEventRecord(c->ev_run1, s));
        c->kevents = false;
    } else {
        HIPCHK(c, hipEventRecord(c->ev_run0, s));
        r = enqueue_kernels(c, a);
        if (r) return r;
        HIPCHK(c, hipEventRecord(c->ev_run1, s));
        c->kevents = true;
    }
    c->last = a;
    c->have_run = true;
    return 0;
}

extern "C" int tess2_run(tess2_ctx *c, int alg, int num_given_bounds, const float *given_mins,
                         const float *given_maxs, int project, const float *proj_plane,
                         float mass, float eps, const int *glo_num_idx, int assignment,
                         int precision, int want_volumes)
{
    if (!c || !glo_num_idx) return fail(c, TESS2_E_ARG, "null argument");
    if (alg != TESS2_DENSE_TESS && alg != TESS2_DENSE_CIC)
        return fail(c, TESS2_E_ARG, "unknown estimator %d", alg);
    if (assignment != TESS2_ASSIGN_REFERENCE_SCAN && assignment != TESS2_ASSIGN_EXACT_VORONOI)
        return fail(c, TESS2_E_ARG, "unknown assignment mode %d", assignment);
    if (precision != TESS2_PREC_MIXED && precision != TESS2_PREC_FP64)
        return fail(c, TESS2_E_ARG, "unknown precision %d", precision);
    if (assignment == TESS2_ASSIGN_EXACT_VORONOI && project)
        return fail(c, TESS2_E_ARG, "the exact assignment has no projection mode");
    if (project) {
        // the reference's index() handles the x-y plane only (src/dense.cpp:1076-1077)
        if (!proj_plane || proj_plane[0] != 0.0f || proj_plane[1] != 0.0f ||
            proj_plane[2] == 0.0f)
            return fail(c, TESS2_E_ARG, "projection supports the x-y plane only");
    }
    HIPCHK(c, hipSetDevice(c->device));
    RunArgs a;
    a.alg = alg;
    a.num_given = num_given_bounds;
    for (int i = 0; i < 3; ++i) {
        a.given_mins[i] = (num_given_bounds > i && given_mins) ? given_mins[i] : 0.0f;
        a.given_maxs[i] = (num_given_bounds > i && given_maxs) ? given_maxs[i] : 0.0f;
        a.gnum[i] = glo_num_idx[i];
    }
    a.project = project ? 1 : 0;
    a.mass = mass;
    a.eps = eps;
    a.assignment = assignment;
    a.precision = precision;
    a.want_vol = want_volumes;
    return enqueue_run(c, a);
}

extern "C" int tess2_sync(tess2_ctx *c, tess2_stats_t *st)
{
    if (!c) return TESS2_E_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    for (int attempt = 0;; ++attempt) {
        HIPCHK(c, hipStreamSynchronize(c->stream));
        if (!c->have_run) return fail(c, TESS2_E_ARG, "no run to wait for");
        bool retry = false;
        for (auto &d : c->blocks) {
            if (d.hm->ctrs[C_ERR] & 4u)
                return fail(c, TESS2_E_CAPACITY,
                            "block %d: a cell exceeds the per-cell capacities", d.gid);
            if (d.hm->ctrs[C_ERR] & 1u) {
                const unsigned want = std::max(2 * d.slow_cap, d.hm->ctrs[C_SLOW] + 1024u);
                int r = dalloc(c, &d.slow, want);
                if (r) return r;
                d.slow_cap = want;
                retry = true;
            }
            if (d.hm->ctrs[C_OVF] > d.ovf_cap) {
                const unsigned long long want =
                    std::max<unsigned long long>(2ull * d.ovf_cap,
                                                 d.hm->ctrs[C_OVF] + d.hm->ctrs[C_OVF] / 4ull);
                if (want > (1ull << 30))
                    return fail(c, TESS2_E_CAPACITY, "block %d: %u shared grid contributions",
                                d.gid, d.hm->ctrs[C_OVF]);
                int r = alloc_lists(c, d, (unsigned)want);
                if (r) return r;
                retry = true;
            }
            if (d.hm->ctrs[C_ERR] & 8u)
                return fail(c, TESS2_E_CAPACITY, "block %d: shared grid point beyond the fold's "
                            "capacities", d.gid);
            if (d.hm->ctrs[C_ERR] & 32u)
                return fail(c, TESS2_E_INTERNAL, "block %d: a shared-list entry's ordinal lies "
                            "outside its point's segment", d.gid);
        }
        if (!retry) break;
        if (attempt >= 4) return fail(c, TESS2_E_CAPACITY, "list capacities did not converge");
        int r = enqueue_run(c, c->last);
        if (r) return r;
    }
    float ms = 0.0f, kms = 0.0f;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev_run0, c->ev_run1));
    c->last_run_ms = ms;
    if (c->kevents) {
        HIPCHK(c, hipEventElapsedTime(&kms, c->ev_k0, c->ev_k1));
        c->ktime_sum += kms;
        c->ktime_n += 1;
    }
    if (st) {
        memset(st, 0, sizeof *st);
        for (auto &d : c->blocks) {
            st->cells_scan += (long long)d.hm->cstat[0];
            st->cells_cic += (long long)d.hm->cstat[1];
            st->cells_incomplete += (long long)d.hm->cstat[2];
            st->cells_outside += (long long)d.hm->cstat[3];
            st->cells_degenerate += (long long)d.hm->cstat[4];
            st->points_offblock += d.hm->ctrs[C_OFFBLOCK];
            st->points_routed += d.hm->ctrs[C_ROUTED];
            st->points_deposited += (long long)d.hm->cstat[5] - d.hm->ctrs[C_OFFBLOCK] -
                                    d.hm->ctrs[C_ROUTED];
            st->points_shared += d.hm->ctrs[C_OVF] + d.hm->ctrs[C_PAIRS];
            st->fp64_fallbacks += d.hm->ctrs[C_FALLBACK] + d.hm->ctrs[C_DEFER];
            st->cells_deferred += d.hm->ctrs[C_SLOW];
            st->tot_mass += d.hm->gsum * (double)d.gp.div;
            float mx;
            memcpy(&mx, &d.hm->gmax, sizeof mx);
            st->max_dense = std::max(st->max_dense, mx);
        }
        st->device_ms = ms;
    }
    return 0;
}

extern "C" int tess2_set_option(tess2_ctx *c, int option, int value)
{
    if (!c) return TESS2_E_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    int *chunk = nullptr;
    switch (option) {
    case TESS2_OPT_GRAPH:
        if (value != 0 && value != 1) return fail(c, TESS2_E_ARG, "graph option is 0 or 1");
        if (!value) drop_graph(c);
        c->graph_on = value != 0;
        return 0;
    case TESS2_OPT_WAVES_PER_CU:
        if (value < 1 || value > 64) return fail(c, TESS2_E_ARG, "waves per CU: 1..64");
        if (value == c->waves_per_cu) return 0;
        c->gen++;
        c->waves_per_cu = value;
        // the chunks interleave the cells over the waves of an XCD
        return c->blocks.size() > 1 ? build_chunks(c) : 0;
    case TESS2_OPT_CELL_KERNEL:
        if (value < -1 || value > 4) return fail(c, TESS2_E_ARG, "cell kernel: -1..4");
        if (value != c->cell_kernel) c->gen++;   // a recorded graph is recorded again
        c->cell_kernel = value;
        return 0;
    case TESS2_OPT_CHUNK_SMALL: chunk = &c->chunk_small; break;
    case TESS2_OPT_CHUNK_MEDIUM: chunk = &c->chunk_medium; break;
    case TESS2_OPT_CHUNK_LARGE: chunk = &c->chunk_large; break;
    default: return fail(c, TESS2_E_ARG, "unknown option %d", option);
    }
    if (value < 1 || value > 4096) return fail(c, TESS2_E_ARG, "chunk size: 1..4096");
    if (*chunk == value) return 0;
    *chunk = value;
    return c->blocks.size() > 1 ? build_chunks(c) : 0;
}

extern "C" int tess2_get_option(tess2_ctx *c, int option, int *value)
{
    if (!c || !value) return TESS2_E_ARG;
    switch (option) {
    case TESS2_OPT_GRAPH: *value = c->graph_on ? 1 : 0; return 0;
    case TESS2_OPT_WAVES_PER_CU: *value = c->waves_per_cu; return 0;
    case TESS2_OPT_CHUNK_SMALL: *value = c->chunk_small; return 0;
    case TESS2_OPT_CHUNK_MEDIUM: *value = c->chunk_medium; return 0;
    case TESS2_OPT_CHUNK_LARGE: *value = c->chunk_large; return 0;
    case TESS2_OPT_CELL_KERNEL: *value = c->cell_kernel; return 0;
    default: return fail(c, TESS2_E_ARG, "unknown option %d", option);
    }
}

// median device time of `reps` runs of the last arguments
static int time_runs(tess2_ctx *c, int reps, float *med)
{
    std::vector<float> t;
    for (int i = 0; i < reps; ++i) {
        int r = enqueue_run(c, c->last);
        if (r) return r;
        r = tess2_sync(c, nullptr);
        if (r) return r;
        t.push_back(c->last_run_ms);
    }
    std::sort(t.begin(), t.end());
    *med = t[t.size() / 2];
    return 0;
}

extern "C" int tess2_autotune(tess2_ctx *c, int reps, float *best_ms)
{
    if (!c) return TESS2_E_ARG;
    if (!c->have_run) return fail(c, TESS2_E_ARG, "autotune needs a previous tess2_run");
    if (reps < 1 || reps > 100) return fail(c, TESS2_E_ARG, "reps: 1..100");
    HIPCHK(c, hipSetDevice(c->device));
    float best = 0.0f;
    int r = time_runs(c, 1, &best);   // warm-up (and capture, in graph mode)
    if (r) return r;
    r = time_runs(c, reps, &best);
    if (r) return r;
    if (c->blocks.size() > 1) {
        // chunk sizes for (small, medium, large) blocks and persistent waves
        static const int chunks[][3] = {{8, 16, 32}, {16, 32, 64}, {32, 64, 128}, {64, 128, 256}};
        static const int waves[] = {8, 16, 24, 32};
        int bc[3] = {c->chunk_small, c->chunk_medium, c->chunk_large}, bw = c->waves_per_cu;
        for (const auto &ch : chunks)
            for (int w : waves) {
                if ((r = tess2_set_option(c, TESS2_OPT_CHUNK_SMALL, ch[0])) ||
                    (r = tess2_set_option(c, TESS2_OPT_CHUNK_MEDIUM, ch[1])) ||
                    (r = tess2_set_option(c, TESS2_OPT_CHUNK_LARGE, ch[2])) ||
                    (r = tess2_set_option(c, TESS2_OPT_WAVES_PER_CU, w)))
                    return r;
                float t = 0.0f;
                if ((r = time_runs(c, 1, &t)) || (r = time_runs(c, reps, &t))) return r;
                if (t < best) {
                    best = t;
                    bc[0] = ch[0];
                    bc[1] = ch[1];
                    bc[2] = ch[2];
                    bw = w;
                }
            }
        if ((r = tess2_set_option(c, TESS2_OPT_CHUNK_SMALL, bc[0])) ||
            (r = tess2_set_option(c, TESS2_OPT_CHUNK_MEDIUM, bc[1])) ||
            (r = tess2_set_option(c, TESS2_OPT_CHUNK_LARGE, bc[2])) ||
            (r = tess2_set_option(c, TESS2_OPT_WAVES_PER_CU, bw)))
            return r;
        // the persistent and the queue-free two-cell kernels with the chosen
        // split, unless a cell kernel was asked for
        if (c->cell_kernel == -1) {
            int bk = -1;
            for (int k : {1, 4}) {
                float t = 0.0f;
                if ((r = tess2_set_option(c, TESS2_OPT_CELL_KERNEL, k)) ||
                    (r = time_runs(c, 1, &t)) || (r = time_runs(c, reps, &t)))
                    return r;
                if (t < best) {
                    best = t;
                    bk = k;
                }
            }
            if ((r = tess2_set_option(c, TESS2_OPT_CELL_KERNEL, bk))) return r;
        }
        float t = 0.0f;
        if ((r = time_runs(c, 1, &t))) return r;   // leave the results of the chosen split
    }
    if (best_ms) *best_ms = best;
    return 0;
}

extern "C" int tess2_kernel_time(tess2_ctx *c, float *avg_ms, int *launches)
{
    if (!c) return TESS2_E_ARG;
    if (avg_ms) *avg_ms = c->ktime_n ? (float)(c->ktime_sum / c->ktime_n) : 0.0f;
    if (launches) *launches = c->ktime_n;
    c->ktime_sum = 0.0;
    c->ktime_n = 0;
    return 0;
}

extern "C" int tess2_fetch(tess2_ctx *c, int b, float *density, float *volumes,
                           float *circumcenters)
{
    if (!c || b < 0 || b >= (int)c->blocks.size()) return fail(c, TESS2_E_ARG, "bad block");
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    DevBlock &d = c->blocks[b];
    if (density && d.density)
        HIPCHK(c, hipMemcpy(density, d.density, sizeof(float) * d.ngrid, hipMemcpyDeviceToHost));
    if (volumes && d.norig) {
        if (!d.vol_valid)
            return fail(c, TESS2_E_ARG, "block %d: the last run computed no volumes", b);
        HIPCHK(c, hipMemcpy(volumes, d.vol, sizeof(float) * d.norig, hipMemcpyDeviceToHost));
    }
    if (circumcenters && d.nt) {
        std::vector<float4> tmp(d.nt);
        HIPCHK(c, hipMemcpy(tmp.data(), d.cc, sizeof(float4) * d.nt, hipMemcpyDeviceToHost));
        for (int t = 0; t < d.nt; ++t) {
            const float4 q = tmp[d.tet_new[t]];   // caller's numbering
            circumcenters[3 * t] = q.x;
            circumcenters[3 * t + 1] = q.y;
            circumcenters[3 * t + 2] = q.z;
        }
    }
    return 0;
}

extern "C" int tess2_fetch_owner(tess2_ctx *c, int b, int *owner)
{
    if (!c || !owner || b < 0 || b >= (int)c->blocks.size())
        return fail(c, TESS2_E_ARG, "bad arguments");
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    DevBlock &d = c->blocks[b];
    if (!d.owner || d.owner_cap < d.ngrid)
        return fail(c, TESS2_E_ARG, "no exact-assignment run on block %d", b);
    HIPCHK(c, hipMemcpy(owner, d.owner, sizeof(int) * d.ngrid, hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int tess2_dense_cells(tess2_ctx *c, int b, int *status, int *npts)
{
    if (!c || b < 0 || b >= (int)c->blocks.size()) return fail(c, TESS2_E_ARG, "bad block");
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    DevBlock &d = c->blocks[b];
    if (status && d.norig)
        HIPCHK(c, hipMemcpy(status, d.status, sizeof(int) * d.norig, hipMemcpyDeviceToHost));
    if (npts && d.norig)
        HIPCHK(c, hipMemcpy(npts, d.npts, sizeof(int) * d.norig, hipMemcpyDeviceToHost));
    return 0;
}

// ---------------------------------------------------------------------------
// one-call entry points
// ---------------------------------------------------------------------------
extern "C" int tess2_fill_circumcenters(tess2_ctx *c, const tess2_tet_t *tets, int num_tets,
                                        const float *particles, int num_particles, float *out)
{
    if (!c || num_tets < 0 || num_particles < 0 || (num_tets && (!tets || !out)))
        return fail(c, TESS2_E_ARG, "bad arguments");
    std::vector<int> v2t(num_particles, -1);
    tess2_block_t b{};
    b.num_orig_particles = 0;
    b.num_particles = num_particles;
    b.particles = particles;
    b.num_tets = num_tets;
    b.tets = tets;
    b.vert_to_tet = v2t.data();
    int r = tess2_stage(c, &b, 1);
    if (r) return r;
    DevBlock &d = c->blocks[0];
    launch_circumcenters(d.pts, d.tets, d.nt, d.cc, c->stream);
    HIPCHK(c, hipGetLastError());
    return tess2_fetch(c, 0, nullptr, nullptr, out);
}

extern "C" int tess2_volumes(tess2_ctx *c, const tess2_block_t *block, float *out)
{
    if (!c || !block || (block->num_orig_particles && !out))
        return fail(c, TESS2_E_ARG, "bad arguments");
    int r = tess2_stage(c, block, 1);
    if (r) return r;
    DevBlock &d = c->blocks[0];
    hipStream_t s = c->stream;
    launch_fill_u32(d.ctrs, C_NCTR, 0u, s);
    launch_circumcenters(d.pts, d.tets, d.nt, d.cc, s);
    BlockIn in = make_in(d);
    CellOut o = make_out(d, true);
    GridP gp{};
    const int ck = c->cell_kernel >= 0 ? c->cell_kernel : 1;
    if (ck == 0) launch_cells_ref(in, gp, o, 0, s);
    else launch_cells_mc(in, nullptr, o, 0, ck, s);   // no density: the grid is not read
    launch_cells_ref_slow(in, gp, o, 0, s);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(d.hm->ctrs, d.ctrs, sizeof(unsigned) * C_NCTR, hipMemcpyDeviceToHost,
                             s));
    HIPCHK(c, hipStreamSynchronize(s));
    if (d.hm->ctrs[C_ERR] & 1u) {
        // more deferred cells than the list holds: one cell per launch slot
        int rr = dalloc(c, &d.slow, (size_t)d.norig);
        if (rr) return rr;
        d.slow_cap = (unsigned)std::max(d.norig, 1);
        o = make_out(d, true);
        launch_fill_u32(d.ctrs, C_NCTR, 0u, s);
        launch_cells_ref(in, gp, o, 0, s);
        launch_cells_ref_slow(in, gp, o, 0, s);
        HIPCHK(c, hipMemcpyAsync(d.hm->ctrs, d.ctrs, sizeof(unsigned) * C_NCTR,
                                 hipMemcpyDeviceToHost, s));
        HIPCHK(c, hipStreamSynchronize(s));
    }
    if (d.hm->ctrs[C_ERR] & 4u)
        return fail(c, TESS2_E_CAPACITY, "a cell exceeds the per-cell capacities");
    d.vol_valid = true;
    return tess2_fetch(c, 0, nullptr, out, nullptr);
}

extern "C" int tess2_volumes_noalloc(tess2_ctx *c, const float *particles, int num_particles,
                                     int num_orig_particles, const tess2_tet_t *tets,
                                     int num_tets, const int *vert_to_tet, float *out)
{
    if (!c) return TESS2_E_ARG;
    const int np = num_particles, norig = num_orig_particles, nt = num_tets;
    if (norig < 0 || np < norig || nt < 0) return fail(c, TESS2_E_ARG, "bad counts");
    if (norig == 0) return 0;
    if (!particles || !tets || !vert_to_tet || !out)
        return fail(c, TESS2_E_ARG, "null array");
    if (reinterpret_cast<uintptr_t>(tets) % 16 != 0)
        return fail(c, TESS2_E_ARG, "tets must be 16-byte aligned");
    HIPCHK(c, hipSetDevice(c->device));
    unsigned char *base = scratch_ptr(c->device);
    if (!base) return fail(c, TESS2_E_HIP, "no scratch address for device %d", c->device);
    // scratch layout: counters, float4 particles, float4 circumcenters,
    // status, point counts, deferred-cell list (one slot per cell, so it
    // cannot overflow); 256-byte aligned regions
    auto up = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t o_ctrs = 0, o_pts = up(o_ctrs + sizeof(unsigned) * C_NCTR),
                 o_cc = up(o_pts + sizeof(float4) * (size_t)np),
                 o_st = up(o_cc + sizeof(float4) * (size_t)nt),
                 o_np = up(o_st + sizeof(int) * (size_t)norig),
                 o_sl = up(o_np + sizeof(int) * (size_t)norig),
                 need = up(o_sl + sizeof(int) * (size_t)norig);
    if (need > NOALLOC_SCRATCH_BYTES)
        return fail(c, TESS2_E_CAPACITY,
                    "%d particles and %d tets need %zu bytes of scratch; the maximum is %zu",
                    np, nt, need, NOALLOC_SCRATCH_BYTES);
    unsigned *ctrs = reinterpret_cast<unsigned *>(base + o_ctrs);
    float4 *pts = reinterpret_cast<float4 *>(base + o_pts);
    float4 *cc = reinterpret_cast<float4 *>(base + o_cc);
    const int4 *t4 = reinterpret_cast<const int4 *>(tets);
    hipStream_t s = c->stream;
    // the counters start at zero on every call (the array keeps the last
    // call's contents); a kernel, not a memset
    launch_fill_u32(ctrs, C_NCTR, 0u, s);
    launch_check_block(t4, nt, vert_to_tet, np, &ctrs[C_ERR], 16u, s);
    unsigned err = 0;
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(&err, &ctrs[C_ERR], sizeof err, hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipStreamSynchronize(s));
    if (err & 16u) return fail(c, TESS2_E_ARG, "a tet or vert_to_tet index is out of range");
    launch_pack_points(particles, np, pts, s);
    launch_circumcenters(pts, t4, nt, cc, s);
    BlockIn in{};
    in.pts = pts;
    in.tets = t4;
    in.v2t = vert_to_tet;
    in.cc = cc;
    in.order = nullptr;
    in.ncells = norig;
    in.ntets = nt;
    CellOut o{};
    o.vol = out;
    o.status = reinterpret_cast<int *>(base + o_st);
    o.npts = reinterpret_cast<int *>(base + o_np);
    o.ctrs = ctrs;
    o.slow = reinterpret_cast<int *>(base + o_sl);
    o.slow_cap = (unsigned)norig;
    o.ncells = norig;
    GridP gp{};
    launch_cells_ref(in, gp, o, 0, s);
    launch_cells_ref_slow(in, gp, o, 0, s);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(&err, &ctrs[C_ERR], sizeof err, hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipStreamSynchronize(s));
    if (err & 4u) return fail(c, TESS2_E_CAPACITY, "a cell exceeds the per-cell capacities");
    return 0;
}

extern "C" int tess2_ghost_sends(tess2_ctx *c, const tess2_block_t *block, int num_neighbors,
                                 const float *neighbor_boxes, int first_neighbor,
                                 unsigned long long *sends)
{
    if (!c || !block || num_neighbors < 0 || num_neighbors > 64 || first_neighbor < 0 ||
        (num_neighbors > 0 && !neighbor_boxes) || (block->num_orig_particles > 0 && !sends))
        return fail(c, TESS2_E_ARG, "bad arguments (at most 64 neighbours)");
    int r = tess2_stage(c, block, 1);
    if (r) return r;
    DevBlock &d = c->blocks[0];
    if (d.norig == 0) return 0;
    float *dbox = nullptr;
    unsigned long long *dsend = nullptr;
    if ((r = dalloc(c, &dbox, (size_t)std::max(6 * num_neighbors, 1))) ||
        (r = dalloc(c, &dsend, (size_t)d.norig)))
        return r;
    hipStream_t s = c->stream;
    if (num_neighbors > 0)
        HIPCHK(c, hipMemcpyAsync(dbox, neighbor_boxes, sizeof(float) * 6 * num_neighbors,
                                 hipMemcpyHostToDevice, s));
    launch_fill_u32(reinterpret_cast<unsigned *>(dsend), 2 * (size_t)d.norig, 0u, s);
    launch_ghost_sends(d.pts, d.tets, d.nt, d.norig, d.bmin, d.bmax, dbox, num_neighbors,
                       first_neighbor, dsend, s);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(sends, dsend, sizeof(unsigned long long) * d.norig,
                             hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipStreamSynchronize(s));
    (void)hipFree(dbox);
    (void)hipFree(dsend);
    return 0;
}

extern "C" int tess2_dense(tess2_ctx *c, int alg, int num_given_bounds, const float *given_mins,
                           const float *given_maxs, int project, const float *proj_plane,
                           float mass, float *data_mins, float *data_maxs, float *grid_phys_mins,
                           float *grid_phys_maxs, float *grid_step_size, float eps,
                           const int *glo_num_idx, const tess2_block_t *blocks, int nblocks,
                           int assignment, int precision, float **density_out,
                           tess2_stats_t *stats)
{
    if (!c || !blocks || nblocks <= 0 || !density_out) return fail(c, TESS2_E_ARG, "bad args");
    int r = tess2_stage(c, blocks, nblocks);
    if (r) return r;
    r = tess2_run(c, alg, num_given_bounds, given_mins, given_maxs, project, proj_plane, mass,
                  eps, glo_num_idx, assignment, precision, 0);
    if (r) return r;
    r = tess2_sync(c, stats);
    if (r) return r;
    for (int i = 0; i < 3; ++i) {
        if (data_mins) data_mins[i] = c->data_mins[i];
        if (data_maxs) data_maxs[i] = c->data_maxs[i];
        if (grid_phys_mins) grid_phys_mins[i] = c->gmin[i];
        if (grid_phys_maxs) grid_phys_maxs[i] = c->gmax[i];
        if (grid_step_size) grid_step_size[i] = c->step[i];
    }
    for (int k = 0; k < nblocks; ++k) {
        if (!density_out[k]) continue;
        r = tess2_fetch(c, k, density_out[k], nullptr, nullptr);
        if (r) return r;
    }
    return 0;
}

extern "C" int tess2_write_grid(const char *path, const float *density, const int *gnum,
                                int project)
{
    if (!path || !density || !gnum) return TESS2_E_ARG;
    FILE *f = fopen(path, "wb");   // truncates, as MPI_File_set_size(fd, 0)
    if (!f) return TESS2_E_IO;
    const size_t n = (size_t)gnum[0] * gnum[1] * (project ? 1 : (size_t)gnum[2]);
    const size_t w = fwrite(density, sizeof(float), n, f);
    const int e = fclose(f);
    return (w == n && e == 0) ? 0 : TESS2_E_IO;
}
