// fold.hip -- the ordered fold of shared grid points.
//
// The reference adds every contribution to a grid point in its own order:
// a block's cells in ascending id (IterateCells, src/dense.cpp:220-313),
// then points received from other blocks (recvd_pts, src/dense.cpp:166-202),
// each as density = (float)((double)density + mass / div)
// (src/dense.cpp:290).  The cell kernels record the first and second
// contribution of a point next to the point and append later ones to a
// shared list; this pass replays the additions of every point that got more
// than one in ascending key order (key = cell id, or the received-point key
// above 2^31):
//   0. pairs: a grid pass adds the points with exactly two contributions;
//   1. gather: each larger point (through its representative list entry,
//      the third contribution) reserves a contiguous segment, sized by its
//      contribution count, is filed in one of three classes by that count,
//      and links its record in head;
//   2. scatter: list entries go to their point's segment, each at the slot
//      its ordinal among the point's contributions gives;
//   3. fold: sort each segment by key and add in order -- in registers (at
//      most 16 contributions, one lane per point), in one wave per point (at
//      most 2048: by lane exchanges up to 64, else in LDS), or by chunk sorts
//      in LDS plus merge passes in global memory (larger, one 1024-lane
//      workgroup per point).
// The additions of one point are a chain of dependent float roundings, so
// each runs on a single lane; only the division by div is done in parallel.
#include <algorithm>

#include "kernels.hpp"

namespace t2 {

__device__ __forceinline__ bool fold_overflowed(const CellOut &o)
{
    return o.ctrs[C_OVF] > o.ovf_cap;   // the host reruns with more room
}

// class list and capacity by a run-time class (selects, not an indexed load:
// an indexed member of the local descriptor would live in scratch)
__device__ __forceinline__ int4 *fold_list(const CellOut &o, int cls)
{
    return cls == 0 ? o.fold[0] : (cls == 1 ? o.fold[1] : o.fold[2]);
}
__device__ __forceinline__ unsigned fold_capacity(const CellOut &o, int cls)
{
    return cls == 0 ? o.fold_cap[0] : (cls == 1 ? o.fold_cap[1] : o.fold_cap[2]);
}

// 0. points with exactly two contributions (the common case: a point on the
// face between two cells) hold both next to the point: added here in the
// order the sort gives the pair (the second first unless its key is
// larger); every point with a second contribution is counted
__global__ __launch_bounds__(256) void k_fold_pairs(const CellOut *__restrict__ outs,
                                                    const GridP *__restrict__ gps)
{
    __shared__ unsigned s_n;
    const CellOut &o = outs[blockIdx.y];
    const double dd = (double)gps[blockIdx.y].div;
    if (threadIdx.x == 0) s_n = 0;
    __syncthreads();
    unsigned np = 0;
    for (unsigned long long g = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
         g < o.ngrid; g += (unsigned long long)gridDim.x * blockDim.x) {
        const unsigned c = o.count[g];
        if (c < 2u) continue;
        ++np;
        if (c != 2u) continue;
        const uint2 f = o.first[g], q = o.second[g];
        const bool sw = q.x > f.x;
        const unsigned m0 = sw ? f.y : q.y, m1 = sw ? q.y : f.y;
        float rho = (float)(0.0 + (double)__uint_as_float(m0) / dd);
        rho = (float)((double)rho + (double)__uint_as_float(m1) / dd);
        o.density[g] = rho;
    }
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) np += __shfl_xor(np, m);
    if ((threadIdx.x & 63) == 0 && np) atomicAdd(&s_n, np);
    __syncthreads();
    if (threadIdx.x == 0 && s_n) atomicAdd(&o.ctrs[C_PAIRS], s_n);
}

// 1. one segment per shared point; slot and segment reservations are
// aggregated per workgroup (one atomic per workgroup and counter per pass of
// 1024 entries: the four counters share a cache line, and a single-block run
// with tens of millions of entries serialises on them)
__global__ __launch_bounds__(1024) void k_fold_gather(const CellOut *__restrict__ outs)
{
    __shared__ unsigned s_cnt[16][4];   // per wave: points of class 0..2, segment total
    const CellOut &o = outs[blockIdx.y];
    if (fold_overflowed(o)) return;
    const unsigned n = o.ctrs[C_OVF];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const unsigned long long below = (1ull << lane) - 1ull;
    for (unsigned base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
        const unsigned e = base + threadIdx.x;
        bool mine = false;
        int lidx = 0, cls = 0;
        unsigned cnt = 0;
        if (e < n) {
            const int4 r = o.ovf[e];
            lidx = r.x;
            mine = r.w == 2;   // the point's third contribution represents it
            if (mine) {
                cnt = o.count[lidx];
                cls = cnt <= FOLD_SMALL_MAX ? 0 : (cnt <= FOLD_MID_MAX ? 1 : 2);
            }
        }
        const unsigned long long m0 = __ballot(mine && cls == 0);
        const unsigned long long m1 = __ballot(mine && cls == 1);
        const unsigned long long m2 = __ballot(mine && cls == 2);
        unsigned incl = cnt;   // inclusive prefix sum of the segment sizes in the wave
        for (int off = 1; off < 64; off <<= 1) {
            const unsigned t = __shfl_up(incl, off);
            if (lane >= off) incl += t;
        }
        if (lane == 63) {
            s_cnt[w][0] = (unsigned)__popcll(m0);
            s_cnt[w][1] = (unsigned)__popcll(m1);
            s_cnt[w][2] = (unsigned)__popcll(m2);
            s_cnt[w][3] = incl;
        }
        __syncthreads();
        if (threadIdx.x < 4) {   // one reservation per counter, then per-wave bases
            const int k = threadIdx.x;
            const int nw = (int)(blockDim.x >> 6);
            unsigned sum = 0;
            for (int i = 0; i < nw; ++i) sum += s_cnt[i][k];
            unsigned b0 = 0;
            if (sum) b0 = atomicAdd(&o.ctrs[k < 3 ? C_FOLD0 + k : C_SEG], sum);
            for (int i = 0; i < nw; ++i) {
                const unsigned c = s_cnt[i][k];
                s_cnt[i][k] = b0;
                b0 += c;
            }
        }
        __syncthreads();
        if (mine) {
            const unsigned long long m = cls == 0 ? m0 : (cls == 1 ? m1 : m2);
            const unsigned idx = s_cnt[w][cls] + (unsigned)__popcll(m & below);
            const unsigned off = s_cnt[w][3] + incl - cnt;
            if (idx >= fold_capacity(o, cls) || (unsigned long long)off + cnt > 3ull * o.ovf_cap ||
                idx >= (1u << 28)) {
                atomicOr(&o.ctrs[C_ERR], 8u);
                o.head[lidx] = -(1 + 3);              // class 3: skipped by the scatter
            } else {
                fold_list(o, cls)[idx] = make_int4(lidx, (int)off, (int)cnt, 0);
                o.seg[off + cnt - 1] = o.first[lidx];
                o.seg[off + cnt - 2] = o.second[lidx];
                o.head[lidx] = -(1 + (int)(idx * 4 + cls));
            }
        }
        __syncthreads();   // s_cnt is reused by the next pass
    }
}

// Passes 2 and 3b deal the items of all blocks (list entries, points) over
// one 1-D grid: the blocks of a multi-block run differ in size by orders of
// magnitude (a cluster core against its outskirts), and a grid per block left
// the largest block's items to its own few workgroups.  Each workgroup builds
// the exclusive prefix of the blocks' item counts in LDS (at most 128 blocks,
// one lane per pair of blocks); an item finds its block by binary search.
enum { FOLD_MAX_BLOCKS = 128 };

// s_pre[b] = items of the blocks below b, s_pre[nb] = all items; items(b)
// is read by the first wave
template <typename F>
__device__ __forceinline__ void fold_block_prefix(unsigned long long *s_pre, int nb, F items)
{
    if (threadIdx.x < 64) {
        const int l = threadIdx.x, b0 = 2 * l, b1 = b0 + 1;
        const unsigned long long c0 = b0 < nb ? items(b0) : 0ull;
        const unsigned long long c1 = b1 < nb ? items(b1) : 0ull;
        unsigned long long incl = c0 + c1;
        for (int off = 1; off < 64; off <<= 1) {
            const unsigned long long t = __shfl_up(incl, off);
            if (l >= off) incl += t;
        }
        const unsigned long long ex = incl - c0 - c1;
        if (b0 <= nb) s_pre[b0] = ex;
        if (b1 <= nb) s_pre[b1] = ex + c0;
        if (b1 + 1 == nb && l == 63) s_pre[nb] = incl;
    }
    __syncthreads();
}

// the block of item e < s_pre[nb]: the largest b with s_pre[b] <= e (an
// empty block shares its start with the next one, so it is never chosen)
__device__ __forceinline__ int fold_block_of(const unsigned long long *s_pre, int nb,
                                             unsigned long long e)
{
    int lo = 0, hi = nb - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (s_pre[mid] <= e) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// 2. list entries into their segments.  Each entry carries its ordinal c
// among its point's contributions (the count the deposit's atomic returned),
// and the first two sit at the segment's end, so the entry's slot is c - 2:
// no atomic, no contention on points with many contributions.  An ordinal
// outside the segment (an internal error) is flagged, not written.
__global__ __launch_bounds__(256) void k_fold_scatter(const CellOut *__restrict__ outs, int nb)
{
    __shared__ unsigned long long s_pre[FOLD_MAX_BLOCKS + 1];
    fold_block_prefix(s_pre, nb, [&](int b) -> unsigned long long {
        const CellOut &o = outs[b];
        return fold_overflowed(o) ? 0ull : (unsigned long long)o.ctrs[C_OVF];
    });
    const unsigned long long n = s_pre[nb];
    for (unsigned long long e = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (unsigned long long)gridDim.x * blockDim.x) {
        const int b = fold_block_of(s_pre, nb, e);
        const CellOut &o = outs[b];
        const int4 r = o.ovf[(unsigned)(e - s_pre[b])];
        const int h = o.head[r.x];
        if (h >= 0) continue;
        const int enc = -h - 1;
        const int cls = enc & 3;
        if (cls == 3) continue;
        const int4 q = fold_list(o, cls)[enc >> 2];   // {point, start, count, 0}
        const unsigned k = (unsigned)r.w - 2u;
        if (k >= (unsigned)q.z - 2u) {
            atomicOr(&o.ctrs[C_ERR], 32u);
            continue;
        }
        o.seg[(unsigned)q.y + k] = make_uint2((unsigned)r.y, (unsigned)r.z);
    }
}

__device__ __forceinline__ void cas_pair(unsigned *kk, unsigned *mm, int j)
{
    const bool sw = kk[j] > kk[j + 1];
    const unsigned k0 = sw ? kk[j + 1] : kk[j], k1 = sw ? kk[j] : kk[j + 1];
    const unsigned m0 = sw ? mm[j + 1] : mm[j], m1 = sw ? mm[j] : mm[j + 1];
    kk[j] = k0;
    kk[j + 1] = k1;
    mm[j] = m0;
    mm[j + 1] = m1;
}

// odd-even transposition sort of the first P register entries (P rounds)
template <int P> __device__ __forceinline__ void oe_sort(unsigned *kk, unsigned *mm)
{
#pragma unroll
    for (int rd = 0; rd < P / 2; ++rd) {
#pragma unroll
        for (int j = 0; j + 1 < P; j += 2) cas_pair(kk, mm, j);
#pragma unroll
        for (int j = 1; j + 1 < P; j += 2) cas_pair(kk, mm, j);
    }
}

// the chain of the first cnt sorted register entries, unrolled over the
// network's size (static register indices, selects rather than branches;
// entries past cnt are skipped, not added as zeros, which would turn a -0
// sum into +0)
template <int P>
__device__ __forceinline__ float reg_chain(const unsigned *mm, int cnt, double dd)
{
    float rho = 0.0f;
#pragma unroll
    for (int j = 0; j < P; ++j) {
        const float t = (float)((double)rho + (double)__uint_as_float(mm[j]) / dd);
        rho = j < cnt ? t : rho;
    }
    return rho;
}

// 3a. at most 16 contributions: one lane per point, odd-even transposition
// sort in registers
__global__ void k_fold_small(const CellOut *__restrict__ outs, const GridP *__restrict__ gps)
{
    const CellOut &o = outs[blockIdx.y];
    const float div = gps[blockIdx.y].div;
    if (fold_overflowed(o)) return;
    unsigned n = o.ctrs[C_FOLD0];
    if (n > o.fold_cap[0]) n = o.fold_cap[0];
    const double dd = (double)div;
    for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += gridDim.x * blockDim.x) {
        const int4 r = o.fold[0][i];
        const uint2 *sg = o.seg + (unsigned)r.y;
        const int cnt = r.z;
        unsigned kk[FOLD_SMALL_MAX], mm[FOLD_SMALL_MAX];
#pragma unroll
        for (int j = 0; j < FOLD_SMALL_MAX; ++j) {
            kk[j] = 0xffffffffu;   // above every key
            mm[j] = 0u;
            if (j < cnt) {
                const uint2 q = sg[j];
                kk[j] = q.x;
                mm[j] = q.y;
            }
        }
        // the network and the chain span the wave's largest count rounded
        // up to 2, 4, 8 or 16 (most shared points have two contributions);
        // the unused entries hold the largest key and stay at the end
        float rho;
        if (__ballot(cnt > 8)) {
            oe_sort<16>(kk, mm);
            rho = reg_chain<16>(mm, cnt, dd);
        } else if (__ballot(cnt > 4)) {
            oe_sort<8>(kk, mm);
            rho = reg_chain<8>(mm, cnt, dd);
        } else if (__ballot(cnt > 2)) {
            oe_sort<4>(kk, mm);
            rho = reg_chain<4>(mm, cnt, dd);
        } else {
            oe_sort<2>(kk, mm);
            rho = reg_chain<2>(mm, cnt, dd);
        }
        o.density[r.x] = rho;
    }
}

// ascending bitonic sort of P (a power of two) {key, value} pairs in LDS by
// the threads of one workgroup
__device__ void lds_bitonic(unsigned *sk, unsigned *sm, int P)
{
    const int nt = blockDim.x;
    for (int k = 2; k <= P; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int t = threadIdx.x; t < P; t += nt) {
                const int u = t ^ j;
                if (u > t) {
                    const unsigned a = sk[t], b = sk[u];
                    if ((a > b) == ((t & k) == 0)) {
                        sk[t] = b;
                        sk[u] = a;
                        const unsigned x = sm[t];
                        sm[t] = sm[u];
                        sm[u] = x;
                    }
                }
            }
            __syncthreads();
        }
    }
}

// after sorting: the quotients mass / div in double, in place (low word in
// sk, high word in sm), computed in parallel
__device__ void lds_quotients(unsigned *sk, unsigned *sm, int m, double dd)
{
    for (int t = threadIdx.x; t < m; t += blockDim.x) {
        const double q = (double)__uint_as_float(sm[t]) / dd;
        const unsigned long long b = (unsigned long long)__double_as_longlong(q);
        sk[t] = (unsigned)b;
        sm[t] = (unsigned)(b >> 32);
    }
    __syncthreads();
}

__device__ __forceinline__ float lds_chain(const unsigned *sk, const unsigned *sm, int m,
                                           float rho)
{
    for (int t = 0; t < m; ++t) {
        const unsigned long long b = ((unsigned long long)sm[t] << 32) | sk[t];
        rho = (float)((double)rho + __longlong_as_double((long long)b));
    }
    return rho;
}

__device__ __forceinline__ int pow2_at_least(int x)
{
    int p = 1;
    while (p < x) p <<= 1;
    return p;
}

// 3b. 17..2048 contributions: one wave per point, bitonic sort by lane
// exchanges (at most 64) or in LDS
__global__ __launch_bounds__(64) void k_fold_mid(const CellOut *__restrict__ outs,
                                                 const GridP *__restrict__ gps, int nb)
{
    __shared__ unsigned sk[FOLD_MID_MAX], sm[FOLD_MID_MAX];
    __shared__ unsigned long long s_pre[FOLD_MAX_BLOCKS + 1];
    fold_block_prefix(s_pre, nb, [&](int b) -> unsigned long long {
        const CellOut &o = outs[b];
        if (fold_overflowed(o)) return 0ull;
        const unsigned n = o.ctrs[C_FOLD0 + 1];
        return n < o.fold_cap[1] ? n : o.fold_cap[1];
    });
    const unsigned long long n = s_pre[nb];
    for (unsigned long long e = blockIdx.x; e < n; e += gridDim.x) {
        const int b = __builtin_amdgcn_readfirstlane(fold_block_of(s_pre, nb, e));
        const CellOut &o = outs[b];
        const double dd = (double)gps[b].div;
        const int4 r = o.fold[1][(unsigned)(e - s_pre[b])];
        const uint2 *sg = o.seg + (unsigned)r.y;
        const int cnt = r.z;
        if (cnt <= 64) {
            // one entry per lane: bitonic network over the wave by lane
            // exchanges, no LDS round trips or barriers; the quotients go
            // to LDS in sorted order for the chain
            const int t = threadIdx.x;
            uint2 q = make_uint2(0xffffffffu, 0u);   // padding sorts last
            if (t < cnt) q = sg[t];
            unsigned k = q.x, m = q.y;
#pragma unroll
            for (int sz = 2; sz <= 64; sz <<= 1) {
#pragma unroll
                for (int j = sz >> 1; j > 0; j >>= 1) {
                    const unsigned pk = __shfl_xor(k, j), pm = __shfl_xor(m, j);
                    // the lower lane of a pair keeps the smaller key in an
                    // ascending run, the larger in a descending one
                    const bool keep_min = ((t & j) == 0) == ((t & sz) == 0);
                    if (keep_min ? pk < k : pk > k) {
                        k = pk;
                        m = pm;
                    }
                }
            }
            const double qd = (double)__uint_as_float(m) / dd;
            const unsigned long long qb = (unsigned long long)__double_as_longlong(qd);
            sk[t] = (unsigned)qb;
            sm[t] = (unsigned)(qb >> 32);
            __syncthreads();
        } else {
            const int P = pow2_at_least(cnt);
            for (int t = threadIdx.x; t < P; t += 64) {
                uint2 q = make_uint2(0xffffffffu, 0u);
                if (t < cnt) q = sg[t];
                sk[t] = q.x;
                sm[t] = q.y;
            }
            __syncthreads();
            lds_bitonic(sk, sm, P);
            lds_quotients(sk, sm, cnt, dd);
        }
        if (threadIdx.x == 0) o.density[r.x] = lds_chain(sk, sm, cnt, 0.0f);
        __syncthreads();
    }
}

// first index i of A in the first k outputs of merge(A, B) (distinct keys)
__device__ __forceinline__ int co_rank(const uint2 *A, int la, const uint2 *B, int lb, int k)
{
    int lo = k > lb ? k - lb : 0, hi = k < la ? k : la;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (A[mid].x < B[k - mid - 1].x) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// 3c. more than 2048 contributions: one 1024-lane workgroup per point;
// chunks of 8192 sorted in LDS, then merge passes between the segment and
// the same range of the (no longer needed) list storage
__global__ __launch_bounds__(1024) void k_fold_large(const CellOut *__restrict__ outs,
                                                     const GridP *__restrict__ gps)
{
    constexpr int CH = 8192;
    __shared__ unsigned sk[CH], sm[CH];
    const CellOut &o = outs[blockIdx.y];
    const float div = gps[blockIdx.y].div;
    if (fold_overflowed(o)) return;
    unsigned n = o.ctrs[C_FOLD0 + 2];
    if (n > o.fold_cap[2]) n = o.fold_cap[2];
    const double dd = (double)div;
    uint2 *tmp = reinterpret_cast<uint2 *>(o.ovf);   // 3 * ovf_cap pairs, like seg
    const int nt = blockDim.x;
    for (unsigned i = blockIdx.x; i < n; i += gridDim.x) {
        const int4 r = o.fold[2][i];
        const int cnt = r.z;
        uint2 *src = o.seg + (unsigned)r.y;
        uint2 *dst = tmp + (unsigned)r.y;
        __syncthreads();   // the previous point is done with the LDS arrays
        for (int c0 = 0; c0 < cnt; c0 += CH) {
            const int m = cnt - c0 < CH ? cnt - c0 : CH;
            const int P = pow2_at_least(m);
            for (int t = threadIdx.x; t < P; t += nt) {
                uint2 q = make_uint2(0xffffffffu, 0u);
                if (t < m) q = src[c0 + t];
                sk[t] = q.x;
                sm[t] = q.y;
            }
            __syncthreads();
            lds_bitonic(sk, sm, P);
            for (int t = threadIdx.x; t < m; t += nt) src[c0 + t] = make_uint2(sk[t], sm[t]);
            __syncthreads();
        }
        for (int w = CH; w < cnt; w <<= 1) {
            for (int s0 = 0; s0 < cnt; s0 += 2 * w) {
                const int la = cnt - s0 < w ? cnt - s0 : w;
                const int rest = cnt - s0 - la;
                const int lb = rest < w ? (rest > 0 ? rest : 0) : w;
                const uint2 *A = src + s0, *B = src + s0 + la;
                const long long L = (long long)la + lb;
                const int k0 = (int)(L * threadIdx.x / nt), k1 = (int)(L * (threadIdx.x + 1) / nt);
                int ia = co_rank(A, la, B, lb, k0), ib = k0 - ia;
                for (int k = k0; k < k1; ++k) {
                    const bool takeA = ib >= lb || (ia < la && A[ia].x < B[ib].x);
                    dst[s0 + k] = takeA ? A[ia++] : B[ib++];
                }
            }
            __syncthreads();
            uint2 *x = src;
            src = dst;
            dst = x;
        }
        float rho = 0.0f;
        for (int c0 = 0; c0 < cnt; c0 += CH) {
            const int m = cnt - c0 < CH ? cnt - c0 : CH;
            for (int t = threadIdx.x; t < m; t += nt) {
                const double q = (double)__uint_as_float(src[c0 + t].y) / dd;
                const unsigned long long b = (unsigned long long)__double_as_longlong(q);
                sk[t] = (unsigned)b;
                sm[t] = (unsigned)(b >> 32);
            }
            __syncthreads();
            if (threadIdx.x == 0) rho = lds_chain(sk, sm, m, rho);
            __syncthreads();
        }
        if (threadIdx.x == 0) o.density[r.x] = rho;
    }
}

static unsigned capped_blocks(unsigned long long work, unsigned per, unsigned cap)
{
    unsigned long long b = (work + per - 1) / per;
    if (b > cap) b = cap;
    if (b == 0) b = 1;
    return (unsigned)b;
}

void launch_fold(const CellOut *outs, const GridP *gps, int nb, unsigned max_ovf_cap,
                 const unsigned max_fold_cap[3], unsigned long long max_grid,
                 unsigned long long sum_ovf_cap, unsigned long long sum_mid_cap, hipStream_t s)
{
    if (nb <= 0) return;
    // per-block grids sized by the largest block's capacities, fewer
    // workgroups per block when there are many blocks (the loops stride);
    // the dealt passes are sized by the capacities of all blocks
    const unsigned per = nb > 16 ? 256 : 4096;
    const unsigned bg = (unsigned)std::min<unsigned long long>(per, max_grid / 256 + 1);
    hipLaunchKernelGGL(k_fold_pairs, dim3(bg, nb), dim3(256), 0, s, outs, gps);
    hipLaunchKernelGGL(k_fold_gather, dim3(capped_blocks(max_ovf_cap, 1024, per / 4), nb),
                       dim3(1024), 0, s, outs);
    hipLaunchKernelGGL(k_fold_scatter, dim3(capped_blocks(sum_ovf_cap, 256, 4096)), dim3(256), 0,
                       s, outs, nb);
    hipLaunchKernelGGL(k_fold_small, dim3(capped_blocks(max_fold_cap[0], 256, per), nb),
                       dim3(256), 0, s, outs, gps);
    hipLaunchKernelGGL(k_fold_mid, dim3(capped_blocks(sum_mid_cap, 1, 8192)), dim3(64), 0, s,
                       outs, gps, nb);
    hipLaunchKernelGGL(k_fold_large, dim3(capped_blocks(max_fold_cap[2], 1, 64), nb),
                       dim3(1024), 0, s, outs, gps);
}

}  // namespace t2
