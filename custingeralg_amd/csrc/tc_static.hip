// tc_static.hip -- static triangle counting on MI355X (gfx950).
//
// Replaces devicecuStingerAllTriangles (src/static_triangle_counting/cct.cu:216-270)
// with the same per-vertex result and a different method:
//
//  * ORIENTED (symmetric, loop-free graphs).  Rank vertices by (degree, id) and
//    split every row into N-(v) (lower rank) and N+(v) (higher rank), each
//    still id-sorted.  A triangle a<b<w (rank order) is found exactly once, by
//    the job that owns b: N+(b) sits in LDS (sorted array + hash of
//    positions), and every task a in N-(b) intersects N+(a) with it.  A hit w
//    adds 1 to an LDS counter of w's slot and to the task's counter, so the
//    per-vertex result 2*t(v) needs at most one global atomic per (job, vertex)
//    instead of one per triangle.
//  * DIRECT (any validated input).  Every entry (a,b) of a row is a task,
//    |N(a) cap N(b)| with the reference's filters (cct.cu:247-249) is added to
//    out[a] only -- the reference's own formula.
//
// Both run on the same kernel: owner list in LDS, tasks flattened across the
// 64 lanes of a wavefront, three strategies per task (lane-sequential merge,
// binary search of the owner list into the task's list, streaming the task's
// list through the LDS hash), chosen by the fixed degree-ratio heuristic or
// the learned table.  Jobs come from a prefix-sum partition of the estimated
// work per owner (the load-balance layer, include/load_balance.cuh:77-102),
// binned by owner-list length (small / mid / large+tiled).
#include "tc_internal.hpp"

#include <algorithm>
#include <cmath>

namespace tc {

enum { S_MERGE = 0, S_BSEARCH = 1, S_HASH = 2, S_MPATH = 3 };
constexpr int kUnroll = 2;
#ifndef TC_BIG_TASK
#define TC_BIG_TASK 32
#endif
#ifndef TC_BIG_TASK_SMALL
#define TC_BIG_TASK_SMALL (1 << 30)
#endif
// hash tasks at least this long are streamed in wave-wide chunks (the small
// bin streams none: its lists are short and the streaming code's registers
// cost its other tasks more than streaming saves; the largest bin, where
// k_count takes hash tasks only under forced policies, walks tasks of 64 or
// more one at a time)
template <int LMAX> constexpr int big_task_min()
{
    return LMAX <= 64 ? TC_BIG_TASK_SMALL : LMAX >= 4096 ? 64 : TC_BIG_TASK;
}
#ifndef TC_BIG_UNROLL
#define TC_BIG_UNROLL 4
#endif
constexpr int kBigUnroll = TC_BIG_UNROLL;  // loads in flight per lane on long tasks

__device__ inline int64_t readlane_i64(int64_t x, int lane)
{
    const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)(uint64_t)x, lane);
    const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)((uint64_t)x >> 32), lane);
    return (int64_t)(((uint64_t)hi << 32) | lo);
}

// Bucketed two-choice hash table in LDS.  Bucket b holds up to S ids (S = 4:
// 16-byte buckets, S = 2: 8-byte buckets) in slots S b .. S b + S - 1, filled
// as a prefix (inserts claim the first empty slot by CAS), -1 when empty.  An
// id goes to its first bucket, to its second when the first is full, and
// when both are full to the next non-full bucket after the first.  A lookup
// reads the first bucket (one LDS read), the second only when the first is
// full, and loops over further buckets only when both are full; the load
// factor <= 1/2 makes each step rarer.
template <int S> struct BucketWord;
template <> struct BucketWord<4> {
    using T = int4;
    __device__ static T empty() { return make_int4(-1, -1, -1, -1); }
    __device__ static int32_t last(const T &v) { return v.w; }
    __device__ static int32_t slot_in(const T &v, int32_t w)
    {
        const int32_t m0 = v.x == w, m1 = v.y == w, m2 = v.z == w, m3 = v.w == w;
        return (m0 | m1 | m2 | m3) ? m1 + 2 * m2 + 3 * m3 : -1;
    }
};
template <> struct BucketWord<2> {
    using T = int2;
    __device__ static T empty() { return make_int2(-1, -1); }
    __device__ static int32_t last(const T &v) { return v.y; }
    __device__ static int32_t slot_in(const T &v, int32_t w)
    {
        const int32_t m0 = v.x == w, m1 = v.y == w;
        return (m0 | m1) ? m1 : -1;
    }
};

template <int S>
struct BucketTableT {
    using W = BucketWord<S>;
    using V = typename W::T;
    static constexpr int kLog = S == 4 ? 2 : 1;
    V *HB;
    int nbb;  // log2(number of buckets)

    __device__ inline uint32_t b1(int32_t w) const
    {
        return nbb ? (((uint32_t)w * 0x9E3779B1u) >> (32 - nbb)) : 0u;
    }
    __device__ inline uint32_t b2(int32_t w) const
    {
        return nbb ? ((((uint32_t)w ^ 0x2C1B3C6Du) * 0x297A2D39u) >> (32 - nbb)) : 0u;
    }
    // A bucket's ids fill its slots in order (claim takes the lowest free
    // one), so a bucket is full exactly when its last slot is taken.
    __device__ inline bool claim(uint32_t b, int32_t w) const
    {
        int *s = reinterpret_cast<int *>(&HB[b]);
#pragma unroll
        for (int k = 0; k < S; ++k)
            if (atomicCAS(&s[k], -1, w) == -1) return true;
        return false;
    }
    // First fit: the first bucket, then the second, then the buckets after
    // the first in turn.  An id is therefore in its first bucket unless that
    // bucket is full, and in its second unless both are (no id leaves the
    // table while it is read).
    __device__ inline void insert(int32_t w) const
    {
        const uint32_t m = (1u << nbb) - 1;
        const uint32_t x = b1(w);
        if (claim(x, w)) return;
        const uint32_t y = b2(w);
        if (claim(y, w)) return;
        for (uint32_t b = (x + 1) & m;; b = (b + 1) & m)
            if (claim(b, w)) return;
    }
    // Slot of w in bucket b or -1; arithmetic only (ids in a bucket are distinct).
    __device__ inline int32_t match(const V v, uint32_t b, int32_t w) const
    {
        const int32_t k = W::slot_in(v, w);
        return k >= 0 ? (int32_t)(b << kLog) + k : -1;
    }
    // The rest of a lookup whose first bucket vx (index x) is full and does
    // not hold w: the second bucket, then the probe sequence if both are full.
    __device__ inline int32_t find_rest(uint32_t x, int32_t w) const
    {
        const uint32_t y = b2(w);
        const V vy = HB[y];
        int32_t s = match(vy, y, w);
        if (s < 0 && W::last(vy) >= 0) {
            const uint32_t m = (1u << nbb) - 1;
            for (uint32_t b = (x + 1) & m;; b = (b + 1) & m) {
                const V v = HB[b];
                s = match(v, b, w);
                if (s >= 0 || W::last(v) < 0) break;
            }
        }
        return s;
    }
    // Slot of w, or -1 (w < 0 is never present).
    __device__ inline int32_t find(int32_t w) const
    {
        const uint32_t x = b1(w);
        const V vx = HB[x];
        int32_t s = match(vx, x, w);
        if (s < 0 && w >= 0 && W::last(vx) >= 0) s = find_rest(x, w);
        return w >= 0 ? s : -1;  // empty slots hold -1
    }
    // Batched find: the first buckets of every value are read before any is
    // matched; a second bucket is read only where the first is full.
    template <int U>
    __device__ inline void find_n(const int32_t *w, int32_t *slot) const
    {
        V vx[U];
        uint32_t x[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            x[u] = b1(w[u]);
            vx[u] = HB[x[u]];
        }
        V vy[U];
        uint32_t y[U];
        int32_t s[U];
        bool need[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            s[u] = match(vx[u], x[u], w[u]);
            need[u] = s[u] < 0 && w[u] >= 0 && W::last(vx[u]) >= 0;
            y[u] = 0;
            vy[u] = W::empty();
            if (need[u]) {  // (these reads are in flight together)
                y[u] = b2(w[u]);
                vy[u] = HB[y[u]];
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (need[u]) {
                s[u] = match(vy[u], y[u], w[u]);
                if (s[u] < 0 && W::last(vy[u]) >= 0) {
                    const uint32_t m = (1u << nbb) - 1;
                    for (uint32_t b = (x[u] + 1) & m;; b = (b + 1) & m) {
                        const V v = HB[b];
                        s[u] = match(v, b, w[u]);
                        if (s[u] >= 0 || W::last(v) < 0) break;
                    }
                }
            }
            slot[u] = w[u] >= 0 ? s[u] : -1;
        }
    }
    __device__ inline int32_t id_at(int32_t slot) const
    {
        return reinterpret_cast<const int32_t *>(HB)[slot];
    }
};
using BucketTable = BucketTableT<4>;

// Membership bitmap of a sorted owner tile over its value range [lo, lo+span):
// word i holds the bits of values lo+32i .. lo+32i+31 (x) and the number of
// set bits in words < i (y), so the position of w in the tile (its counter
// slot) is y plus the set bits below w in x.  One 8-byte LDS read per probe.
// Offsets are clamped to guard = 32 * (number of words): values outside the
// range (w < lo and w = -1 wrap high) read the zero word after the last one,
// so a probe needs no range test.  find_n probes a batch branch-free: all
// reads are issued before any is used.
struct TileBitmap {
    const uint2 *BM;
    int32_t lo;
    uint32_t guard;

    template <int U>
    __device__ inline void find_n(const int32_t (&w)[U], int32_t (&slot)[U]) const
    {
        uint2 x[U];
        uint32_t off[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            off[u] = min((uint32_t)(w[u] - lo), guard);
            x[u] = BM[off[u] >> 5];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            // v_bfe_u32 reads bits 4:0 of its offset and width
            const bool hit = __builtin_amdgcn_ubfe(x[u].x, off[u], 1u) != 0u;
            const int32_t pos = (int32_t)(x[u].y + __popc(__builtin_amdgcn_ubfe(x[u].x, 0u, off[u])));
            slot[u] = hit ? pos : -1;
        }
    }
};

// ------------------------------------------------------------ helpers --

__device__ inline void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ inline int32_t wave_incl_scan_i32(int32_t x)
{
    const int lane = threadIdx.x & (kWave - 1);
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        int32_t y = __shfl_up(x, d, kWave);
        if (lane >= d) x += y;
    }
    return x;
}

__host__ __device__ inline int ceil_log2_i(int64_t x)
{
    int r = 0;
    while ((int64_t(1) << r) < x) ++r;
    return r;
}

// Largest r in [lo,hi] with coff[r] <= e.
__device__ inline int32_t row_search(const int64_t *__restrict__ coff, int64_t e, int32_t lo,
                                     int32_t hi)
{
    while (lo < hi) {
        int32_t mid = lo + ((hi - lo + 1) >> 1);
        if (coff[mid] <= e)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// Row of each of the 64 entries [64c, 64c+64) of a wave.  crow[c] is the row
// holding entry 64c (k_chunk_rows); the rows of the chunk lie in
// [crow[c], crow[c+1]].  Entries past nnz get the last row; callers mask them.
__device__ inline int32_t wave_row_of(const int64_t *__restrict__ coff,
                                      const int32_t *__restrict__ crow, int32_t nv, int64_t c,
                                      int64_t nnz)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t e0 = c * kWave;
    int64_t e = e0 + lane;
    const int64_t elast = min(e0 + kWave, nnz) - 1;
    const int64_t nchunks = ceil_div(nnz, kWave);
    const int32_t rfirst = crow[c];
    const int32_t rlast = (c + 1 < nchunks) ? crow[c + 1] : nv - 1;
    if (e > elast) e = elast;
    return row_search(coff, e, rfirst, rlast);
}

// wave_row_of for two chunks c1 and c2, their binary searches in lockstep so
// that the loads of both are in flight together.
__device__ inline void wave_rows_of2(const int64_t *__restrict__ coff,
                                     const int32_t *__restrict__ crow, int32_t nv, int64_t c1,
                                     int64_t c2, int64_t nnz, int32_t &r1, int32_t &r2)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t nchunks = ceil_div(nnz, kWave);
    const int64_t e1 = min(c1 * kWave + lane, min(c1 * kWave + kWave, nnz) - 1);
    const int64_t e2 = min(c2 * kWave + lane, min(c2 * kWave + kWave, nnz) - 1);
    int32_t lo1 = crow[c1], hi1 = (c1 + 1 < nchunks) ? crow[c1 + 1] : nv - 1;
    int32_t lo2 = crow[c2], hi2 = (c2 + 1 < nchunks) ? crow[c2 + 1] : nv - 1;
    while (lo1 < hi1 || lo2 < hi2) {
        const int32_t m1 = lo1 + ((hi1 - lo1 + 1) >> 1), m2 = lo2 + ((hi2 - lo2 + 1) >> 1);
        const int64_t v1 = lo1 < hi1 ? coff[m1] : 0, v2 = lo2 < hi2 ? coff[m2] : 0;
        if (lo1 < hi1) {
            if (v1 <= e1) lo1 = m1; else hi1 = m1 - 1;
        }
        if (lo2 < hi2) {
            if (v2 <= e2) lo2 = m2; else hi2 = m2 - 1;
        }
    }
    r1 = lo1;
    r2 = lo2;
}

// crow[c] = row holding compact entry 64c.  A row holds the starts of chunks
// ceil(a/64) .. ceil(b/64)-1 ([a, b) its entries); a row with more than 8 of
// them is written by its whole wave.
__global__ void k_chunk_rows(const int64_t *__restrict__ coff, int32_t nv, int32_t *__restrict__ crow)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t r0 = (int64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~(kWave - 1)); r0 < nv;
         r0 += stride) {
        const int64_t r = r0 + lane;
        int64_t c0 = 0, c1 = 0;
        if (r < nv) {
            c0 = ceil_div(coff[r], kWave);
            c1 = ceil_div(coff[r + 1], kWave);
        }
        const bool wide = c1 - c0 > 8;
        if (!wide)
            for (int64_t c = c0; c < c1; ++c) crow[c] = (int32_t)r;
        for (uint64_t wm = __ballot(wide); wm; wm &= wm - 1) {
            const int j = __ffsll((unsigned long long)wm) - 1;
            const int64_t d0 = readlane_i64(c0, j), d1 = readlane_i64(c1, j);
            for (int64_t c = d0 + lane; c < d1; c += kWave) crow[c] = (int32_t)(r0 + j);
        }
    }
}

__device__ inline bool rank_less(int32_t a, int32_t la, int32_t b, int32_t lb)
{
    return (la < lb) || (la == lb && a < b);
}

__global__ void k_fill_u32(uint32_t *p, int64_t n, uint32_t v)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// Offsets of the five row lists of the N+ sorts, and their counters zeroed.
struct ListOffsets {
    int64_t v[5];
};
__global__ void k_set_list_offsets(int64_t *loff, ListOffsets o, uint32_t *lcnt)
{
    if (threadIdx.x < 5) loff[threadIdx.x] = o.v[threadIdx.x];
    if (threadIdx.x < 8) lcnt[threadIdx.x] = 0u;
}

__global__ void k_fill_u64(unsigned long long *p, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0ull;
}

// ---------------------------------------------------------- orientation --

// mask[c] = bit i set iff compact entry 64c+i points to a higher-rank vertex.
__global__ void k_orient_mask(const int64_t *__restrict__ coff, const int32_t *__restrict__ crow,
                              const int64_t *__restrict__ boff,
                              const int32_t *__restrict__ len, const int32_t *__restrict__ pool,
                              int32_t nv, int64_t nnz, uint64_t *__restrict__ mask,
                              int32_t *__restrict__ cnt)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t nchunks = ceil_div(nnz, kWave);
    const int64_t wstride = (int64_t)gridDim.x * (blockDim.x / kWave);
    for (int64_t c = (int64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; c < nchunks;
         c += wstride) {
        int64_t e = c * kWave + lane;
        int32_t a = wave_row_of(coff, crow, nv, c, nnz);
        bool plus = false;
        if (e < nnz) {
            int32_t b = pool[boff[a] + (e - coff[a])];
            plus = rank_less(a, len[a], b, len[b]);
        }
        uint64_t m = __ballot(plus);
        if (lane == 0) {
            mask[c] = m;
            cnt[c] = __popcll(m);
        }
    }
}

__device__ inline int64_t plus_prefix(const int64_t *__restrict__ cscan,
                                      const uint64_t *__restrict__ mask, int64_t x)
{
    int64_t c = x >> 6;
    int r = (int)(x & 63);
    uint64_t m = r ? (mask[c] & (~0ull >> (64 - r))) : 0ull;
    return cscan[c] + __popcll(m);
}

// dmo[v] = |N-(v)| of vertex v (store order).
__global__ void k_orient_dm(const int64_t *__restrict__ coff, const int32_t *__restrict__ len,
                            const int64_t *__restrict__ cscan, const uint64_t *__restrict__ mask,
                            int32_t nv, int32_t *__restrict__ dmo)
{
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < nv;
         v += (int64_t)gridDim.x * blockDim.x) {
        const int64_t p0 = plus_prefix(cscan, mask, coff[v]);
        const int64_t p1 = plus_prefix(cscan, mask, coff[v + 1]);
        dmo[v] = len[v] - (int32_t)(p1 - p0);
    }
}

// Sort keys of the rank order: the degree, with the vertex id as payload (the
// LSD radix sort is stable, so ties stay in id order: rank = (degree, id)).
// sc[0] (zero on entry) receives the maximum degree, which bounds the key bits.
__global__ void k_rank_keys(const int32_t *__restrict__ len, int32_t nv, uint64_t *__restrict__ keys,
                            int32_t *__restrict__ ids, unsigned long long *__restrict__ sc)
{
    __shared__ unsigned long long part[16];
    unsigned long long m = 0;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < nv;
         v += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t d = (uint32_t)len[v];
        keys[v] = (uint64_t)d;
        ids[v] = (int32_t)v;
        m = max(m, (unsigned long long)d);
    }
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) m = max(m, (unsigned long long)__shfl_xor(m, d, kWave));
    if ((threadIdx.x & (kWave - 1)) == 0) part[threadIdx.x / kWave] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < (int)(blockDim.x / kWave); ++w) m = max(m, part[w]);
        if (m) atomicMax(&sc[0], m);
    }
}

// Batched graphs: keys |= graph of the vertex << dbits, so ranks run graph by
// graph ((graph, degree, id) order).  Every edge lies inside one graph, where
// the order is still (degree, id): the orientation is unchanged, and each
// graph's plus lists and task records become contiguous.
__global__ void k_rank_keys_graph(const int64_t *__restrict__ voff, int32_t G, int32_t nv, int dbits,
                                  uint64_t *__restrict__ keys)
{
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < nv;
         v += (int64_t)gridDim.x * blockDim.x) {
        int32_t lo = 0, hi = G - 1;
        while (lo < hi) {
            const int32_t mid = (lo + hi + 1) >> 1;
            if (voff[mid] <= v) lo = mid; else hi = mid - 1;
        }
        keys[v] |= (uint64_t)lo << dbits;
    }
}

// rnk[order[r]] = r.
__global__ void k_rank_inverse(const int32_t *__restrict__ order, int32_t nv, int32_t *__restrict__ rnk)
{
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < nv;
         r += (int64_t)gridDim.x * blockDim.x)
        rnk[order[r]] = (int32_t)r;
}

// rnk[order[r]] = r, and |N+| / |N-| of the vertex of rank r.
__global__ void k_rank_of(const int32_t *__restrict__ order, const int32_t *__restrict__ len,
                          const int32_t *__restrict__ dmo, int32_t nv, int32_t *__restrict__ rnk,
                          int32_t *__restrict__ dpr, int32_t *__restrict__ dmr)
{
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < nv;
         r += (int64_t)gridDim.x * blockDim.x) {
        const int32_t v = order[r];
        rnk[v] = (int32_t)r;
        dpr[r] = len[v] - dmo[v];
        dmr[r] = dmo[v];
    }
}

// Rank-space views: N+ of rank r starts at pcoff[r] of the plus array.
__global__ void k_rank_vinfo(const int64_t *__restrict__ pcoff, const int32_t *__restrict__ dpr,
                             const int32_t *__restrict__ dmr, int32_t nv, VInfo *__restrict__ vi)
{
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < nv;
         r += (int64_t)gridDim.x * blockDim.x) {
        VInfo x;
        x.rb = pcoff[r];
        x.dm = dmr[r];
        x.len = dpr[r] + dmr[r];
        vi[r] = x;
    }
}

// Writes N+ of every vertex into the plus array of the rank layout, as ranks
// (unsorted yet); the task array is filled by k_orient_transpose.
__global__ void k_orient_write_plus(const int64_t *__restrict__ coff, const int32_t *__restrict__ crow,
                                    const int64_t *__restrict__ boff, const int32_t *__restrict__ pool,
                                    const int64_t *__restrict__ cscan,
                                    const uint64_t *__restrict__ mask,
                                    const int32_t *__restrict__ rnk, const int64_t *__restrict__ pcoff,
                                    int32_t nv, int64_t nnz, int32_t *__restrict__ padj)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t nchunks = ceil_div(nnz, kWave);
    const int64_t wstride = (int64_t)gridDim.x * (blockDim.x / kWave);
    const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
    for (int64_t c = (int64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; c < nchunks;
         c += wstride) {
        const int64_t e = c * kWave + lane;
        const int32_t a = wave_row_of(coff, crow, nv, c, nnz);
        if (e >= nnz) continue;
        const uint64_t m = mask[c];
        if (!((m >> lane) & 1ull)) continue;
        const int64_t pe = cscan[c] + __popcll(m & lt);
        const int64_t ca = coff[a];
        const int64_t rp = pe - plus_prefix(cscan, mask, ca);
        const int32_t b = pool[boff[a] + (e - ca)];
        padj[pcoff[rnk[a]] + rp] = rnk[b];
    }
}

// With the ranks known first: the orientation mask by rank (rank(b) >
// rank(a) is (degree, id) order, or (graph, degree, id) for batched graphs,
// whose edges stay inside one graph) and every entry's neighbour as a rank,
// tr[e], so that the N+ write reads those in order instead of looking up the
// rank of each neighbour again.
__global__ void k_orient_mask_rank(const int64_t *__restrict__ coff, const int32_t *__restrict__ crow,
                                   const int64_t *__restrict__ boff, const int32_t *__restrict__ pool,
                                   const int32_t *__restrict__ rnk, int32_t nv, int64_t nnz,
                                   uint64_t *__restrict__ mask, int32_t *__restrict__ cnt,
                                   int32_t *__restrict__ tr)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t nchunks = ceil_div(nnz, kWave);
    const int64_t wstride = (int64_t)gridDim.x * (blockDim.x / kWave);
    const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
    // two chunks per pass (c and c + wstride): their row searches in
    // lockstep, then both chunks' loads and rank lookups in flight together
    for (int64_t c = (int64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; c < nchunks;
         c += 2 * wstride) {
        const int64_t c2 = c + wstride;
        const bool two = c2 < nchunks;  // (wave-uniform)
        const int64_t e = c * kWave + lane, e2 = c2 * kWave + lane;
        int32_t a, a2;
        wave_rows_of2(coff, crow, nv, c, two ? c2 : c, nnz, a, a2);
        const bool in = e < nnz, in2 = two && e2 < nnz;
        const int32_t b = in ? pool[boff[a] + (e - coff[a])] : 0;
        const int32_t b2 = in2 ? pool[boff[a2] + (e2 - coff[a2])] : 0;
        const int32_t ra = in ? rnk[a] : 0, ra2 = in2 ? rnk[a2] : 0;
        const int32_t rb = in ? rnk[b] : 0, rb2 = in2 ? rnk[b2] : 0;
        const bool plus = in && rb > ra, plus2 = in2 && rb2 > ra2;
        const uint64_t m = __ballot(plus), m2 = __ballot(plus2);
        // each chunk's plus entries, compacted to the front of its 64 slots
        if (plus) tr[c * kWave + __popcll(m & lt)] = rb;
        if (plus2) tr[c2 * kWave + __popcll(m2 & lt)] = rb2;
        if (lane == 0) {
            mask[c] = m;
            cnt[c] = __popcll(m);
            if (two) {
                mask[c2] = m2;
                cnt[c2] = __popcll(m2);
            }
        }
    }
}

// Per store row a: where its plus entries go, less their store-order rank
// among all plus entries (pdst[a] + plus index = slot in the plus array).
__global__ void k_plus_dst(const int64_t *__restrict__ coff, const int64_t *__restrict__ cscan,
                           const uint64_t *__restrict__ mask, const int32_t *__restrict__ rnk,
                           const int64_t *__restrict__ pcoff, int32_t nv, int64_t *__restrict__ pdst)
{
    for (int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; a < nv;
         a += (int64_t)gridDim.x * blockDim.x)
        pdst[a] = pcoff[rnk[a]] - plus_prefix(cscan, mask, coff[a]);
}

// k_orient_write_plus from the entries' ranks (k_orient_mask_rank, compacted
// per chunk) and the rows' destinations (k_plus_dst).
__global__ void k_orient_write_plus_rank(const int64_t *__restrict__ coff,
                                         const int32_t *__restrict__ crow,
                                         const int32_t *__restrict__ tr,
                                         const int64_t *__restrict__ cscan,
                                         const uint64_t *__restrict__ mask,
                                         const int64_t *__restrict__ pdst, int32_t nv, int64_t nnz,
                                         int32_t *__restrict__ padj)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t nchunks = ceil_div(nnz, kWave);
    const int64_t wstride = (int64_t)gridDim.x * (blockDim.x / kWave);
    const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
    // two chunks per pass (c and c + wstride), their row searches in
    // lockstep; the entries' ranks and the chunks' offsets are in flight
    // during the searches
    for (int64_t c = (int64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; c < nchunks;
         c += 2 * wstride) {
        const int64_t c2 = c + wstride;
        const bool two = c2 < nchunks;  // (wave-uniform)
        const int64_t e = c * kWave + lane, e2 = c2 * kWave + lane;
        const uint64_t m = mask[c], m2 = two ? mask[c2] : 0ull;
        if (!m && !m2) continue;  // (the whole wave)
        const bool mine = e < nnz && ((m >> lane) & 1ull);
        const bool mine2 = e2 < nnz && ((m2 >> lane) & 1ull);
        const int32_t k = __popcll(m & lt), k2 = __popcll(m2 & lt);
        const int32_t rb = mine ? tr[c * kWave + k] : 0;
        const int32_t rb2 = mine2 ? tr[c2 * kWave + k2] : 0;
        const int64_t cs = m ? cscan[c] : 0, cs2 = m2 ? cscan[c2] : 0;
        int32_t a, a2;
        wave_rows_of2(coff, crow, nv, m ? c : c2, m2 ? c2 : c, nnz, a, a2);
        if (mine) padj[pdst[a] + cs + k] = rb;
        if (mine2) padj[pdst[a2] + cs2 + k2] = rb2;
    }
}

// One lane per rank row: N+ parts of 2..kThreadSort entries are sorted in
// registers by a bitonic network (of 16 entries where a lane of the wave has
// more than 8, else of 8), longer ones are listed by size class:
// (kThreadSort,64] -> lst[0], (64,512] -> lst[1], (512,4096] -> lst[2],
// (4096,16384] -> lst[3], above -> lst[4].  List q starts at lst + off[q],
// its length is cnt[q].
#ifndef TC_THREAD_SORT
#define TC_THREAD_SORT 16
#endif
constexpr int kThreadSort = TC_THREAD_SORT;  // 8 or 16
template <int N>
__device__ inline void bitonic_regs(int32_t (&v)[N])
{
#pragma unroll
    for (int k = 2; k <= N; k <<= 1) {
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
#pragma unroll
            for (int i = 0; i < N; ++i) {
                const int l = i ^ j;
                if (l > i) {
                    const bool up = (i & k) == 0;
                    const int32_t a = v[i], b = v[l];
                    v[i] = up ? min(a, b) : max(a, b);
                    v[l] = up ? max(a, b) : min(a, b);
                }
            }
        }
    }
}
__global__ void k_sort_plus_small(const VInfo *__restrict__ vi, int32_t nv,
                                  int32_t *__restrict__ padj, int32_t *__restrict__ lst,
                                  const int64_t *__restrict__ off /* [5] */,
                                  int32_t *__restrict__ cnt /* [5] */)
{
    const int lane = threadIdx.x & (kWave - 1);
    const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
    // whole waves iterate together so the list appends can be aggregated
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t r0 = (int64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~(kWave - 1)); r0 < nv;
         r0 += stride) {
        const int64_t r = r0 + lane;
        int32_t dp = 0;
        VInfo x = {0, 0, 0};
        if (r < nv) {
            x = vi[r];
            dp = x.len - x.dm;
        }
        const int q = dp <= kThreadSort ? -1
                              : (dp <= 64 ? 0 : (dp <= 512 ? 1 : (dp <= 4096 ? 2 : (dp <= 16384 ? 3 : 4))));
#pragma unroll
        for (int c = 0; c < 5; ++c) {
            const uint64_t m = __ballot(q == c);
            if (!m) continue;
            int32_t base = 0;
            if (lane == __ffsll((unsigned long long)m) - 1) base = atomicAdd(&cnt[c], __popcll(m));
            base = __shfl(base, __ffsll((unsigned long long)m) - 1, kWave);
            if (q == c) lst[off[c] + base + __popcll(m & lt)] = (int32_t)r;
        }
        const bool mine = dp >= 2 && dp <= kThreadSort;
        int32_t *P = padj + x.rb;
        if (kThreadSort > 8 && __ballot(mine && dp > 8)) {
            int32_t v[kThreadSort > 8 ? kThreadSort : 8];
#pragma unroll
            for (int i = 0; i < kThreadSort; ++i) v[i] = (mine && i < dp) ? P[i] : INT32_MAX;
            bitonic_regs(v);
#pragma unroll
            for (int i = 0; i < kThreadSort; ++i)
                if (mine && i < dp) P[i] = v[i];
        } else if (__ballot(mine)) {
            int32_t v[8];
#pragma unroll
            for (int i = 0; i < 8; ++i) v[i] = (mine && i < dp) ? P[i] : INT32_MAX;
            bitonic_regs(v);
#pragma unroll
            for (int i = 0; i < 8; ++i)
                if (mine && i < dp) P[i] = v[i];
        }
    }
}

// N+ parts of kThreadSort+1..64 entries: one wavefront per listed row,
// bitonic network across the lanes.
__global__ void k_sort_plus_wave(const VInfo *__restrict__ vi, const int32_t *__restrict__ rows,
                                 const int32_t *__restrict__ nrows, int32_t *__restrict__ padj)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int32_t n = *nrows;
    const int64_t wstride = (int64_t)gridDim.x * (blockDim.x / kWave);
    for (int64_t q = (int64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; q < n;
         q += wstride) {
        const VInfo x = vi[rows[q]];
        const int32_t dp = x.len - x.dm;
        int32_t *P = padj + x.rb;
        int32_t v = lane < dp ? P[lane] : INT32_MAX;
#pragma unroll
        for (int k = 2; k <= kWave; k <<= 1) {
#pragma unroll
            for (int j = k >> 1; j > 0; j >>= 1) {
                const int32_t y = __shfl_xor(v, j, kWave);
                const bool keep_min = ((lane & j) == 0) == ((lane & k) == 0);
                v = keep_min ? min(v, y) : max(v, y);
            }
        }
        if (lane < dp) P[lane] = v;
    }
}

// Sort of the N+ part of the listed rows (one block per row, rows of at most
// CAP entries): each thread sorts a run of PER = CAP / NT entries in
// registers, then runs are merged pairwise in LDS, each thread producing PER
// outputs of its pair after a merge-path split (a binary search on the
// diagonal).  About PER + log2(run) LDS reads per thread and level, against
// PER * log2(CAP) per level of a bitonic network over the whole row.  Only
// the power of two of runs that covers the row takes part.
#ifndef TC_SORT512_BLOCKS
#define TC_SORT512_BLOCKS 64  // blocks per CU of the 512-entry row sort
#endif
#ifndef TC_SORT4096_BLOCKS
#define TC_SORT4096_BLOCKS 8  // blocks per CU of the 4096-entry row sort
#endif
template <int CAP, int NT>
__global__ __launch_bounds__(NT) void k_sort_plus_block(const VInfo *__restrict__ vi,
                                                        const int32_t *__restrict__ rows,
                                                        const int32_t *__restrict__ nrows,
                                                        int32_t *__restrict__ padj)
{
    constexpr int PER = CAP / NT;
    static_assert(PER * NT == CAP && (PER & (PER - 1)) == 0, "runs of a power of two");
    __shared__ int32_t S[CAP];
    const int t = threadIdx.x;
    const int32_t n = *nrows;
    for (int32_t q = blockIdx.x; q < n; q += gridDim.x) {
        const VInfo x = vi[rows[q]];
        const int32_t dp = x.len - x.dm;
        int32_t *P = padj + x.rb;
        const int32_t runs = (dp + PER - 1) / PER;
        int32_t m2 = 1;
        while (m2 < runs) m2 <<= 1;
        const int32_t m = m2 * PER;
        const bool act = t < m2;
        for (int32_t i = t; i < m; i += NT) S[i] = i < dp ? P[i] : INT32_MAX;
        __syncthreads();
        int32_t r[PER];
        if (act) {
#pragma unroll
            for (int k = 0; k < PER; ++k) r[k] = S[t * PER + k];
#pragma unroll
            for (int k = 2; k <= PER; k <<= 1) {
#pragma unroll
                for (int j = k >> 1; j > 0; j >>= 1) {
#pragma unroll
                    for (int i = 0; i < PER; ++i) {
                        const int l = i ^ j;
                        if (l > i) {
                            const int32_t lo = min(r[i], r[l]), hi = max(r[i], r[l]);
                            const bool up = (i & k) == 0;
                            r[i] = up ? lo : hi;
                            r[l] = up ? hi : lo;
                        }
                    }
                }
            }
#pragma unroll
            for (int k = 0; k < PER; ++k) S[t * PER + k] = r[k];  // its own run only
        }
        __syncthreads();
        for (int32_t w = PER; w < m; w <<= 1) {
            if (act) {
                const int32_t d0 = t * PER, base = d0 & ~(2 * w - 1), d = d0 - base;
                const int32_t *A = S + base, *B = S + base + w;
                int32_t lo = max(0, d - w), hi = min(d, w);
                while (lo < hi) {  // entries taken from A among the first d
                    const int32_t mid = (lo + hi) >> 1;
                    if (A[mid] <= B[d - 1 - mid]) lo = mid + 1; else hi = mid;
                }
                int32_t i = lo, j = d - lo;
#pragma unroll
                for (int k = 0; k < PER; ++k) {
                    const int32_t av = i < w ? A[i] : INT32_MAX, bv = j < w ? B[j] : INT32_MAX;
                    const bool ta = j >= w || (i < w && av <= bv);
                    r[k] = ta ? av : bv;
                    i += ta ? 1 : 0;
                    j += ta ? 0 : 1;
                }
            }
            __syncthreads();  // every read of this level is done
            if (act) {
#pragma unroll
                for (int k = 0; k < PER; ++k) S[t * PER + k] = r[k];
            }
            __syncthreads();
        }
        for (int32_t i = t; i < dp; i += NT) P[i] = S[i];
        __syncthreads();
    }
}

__global__ void k_widen_u32(const int32_t *__restrict__ x, int64_t n, uint64_t *__restrict__ y)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        y[i] = (uint64_t)(uint32_t)x[i];
}

__global__ void k_narrow_u64(const uint64_t *__restrict__ y, int64_t n, int32_t *__restrict__ x)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        x[i] = (int32_t)y[i];
}

// Fills the task array: every N+ entry (rank row r, position k, value w)
// becomes the task (r, |N+(r)| - k - 1) of owner w: the task vertex and the
// length of the suffix of N+(r) after w, the only part that can hit.  The
// order of an owner's tasks is arbitrary.
__global__ void k_orient_transpose(const int64_t *__restrict__ pcoff, const int32_t *__restrict__ pcrow,
                                   const int32_t *__restrict__ padj,
                                   const int64_t *__restrict__ toff, int32_t nv, int64_t nplus,
                                   TaskRec *__restrict__ tasks, int32_t *__restrict__ cursor,
                                   unsigned long long *__restrict__ sc, int32_t *__restrict__ bad)
{
    // owners from rank sc[1] up are the hubs, handled by k_orient_transpose_hubs
    const int32_t hub_lo = (int32_t)sc[1];
    unsigned long long nhub = 0;
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t nchunks = ceil_div(nplus, kWave);
    const int64_t wstride = (int64_t)gridDim.x * (blockDim.x / kWave);
    for (int64_t c = (int64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; c < nchunks;
         c += wstride) {
        const int64_t e = c * kWave + lane;
        const int32_t r = wave_row_of(pcoff, pcrow, nv, c, nplus);
        if (e >= nplus) continue;
        const int32_t w = padj[e];
        if (w >= hub_lo) {
            ++nhub;  // filled by k_orient_transpose_hubs
            continue;
        }
        const int32_t slot = atomicAdd(&cursor[w], 1);
        // |N-(w)| from the task offsets (one cache line with toff[w]) rather
        // than from vi[w], a third random read per entry
        const int64_t t0 = toff[w];
        if (slot < toff[w + 1] - t0)
            tasks[t0 + slot] = TaskRec::make(e + 1, (int32_t)(pcoff[r + 1] - e - 1));
        else
            atomicOr(bad, 1);  // rows are not symmetric
    }
    const unsigned long long part[1] = {nhub};
    block_add_u64<1>(part, sc + 2);
}

// Last rank r in [lo, hi] whose rows start at or before x in the degree
// prefix pcoff[r] + toff[r] (= sum of the degrees of the ranks below r).
__device__ inline int32_t hub_of(const int64_t *__restrict__ pcoff, const int64_t *__restrict__ toff,
                                 int64_t x, int32_t lo, int32_t hi)
{
    while (lo < hi) {
        const int32_t mid = lo + ((hi - lo + 1) >> 1);
        if (pcoff[mid] + toff[mid] <= x) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Task lists of the hubs (ranks from sc[1] up), built from the hubs' own store
// rows instead of by appends: the k-th entry of the row of hub w that is
// ranked below w is its task number k - (plus entries before it), and the
// task's suffix starts after w's position in N+(a), found by binary search.
// A hub gets an append from every part of the plus array, so its cursor would
// serialise up to deg(w) atomics on one address.
__global__ void k_orient_transpose_hubs(
    const int64_t *__restrict__ pcoff, const int64_t *__restrict__ toff,
    const int32_t *__restrict__ order, const int32_t *__restrict__ rnk,
    const int64_t *__restrict__ ocoff, const int64_t *__restrict__ boff,
    const int32_t *__restrict__ pool, const uint64_t *__restrict__ mask,
    const int64_t *__restrict__ cscan, const int32_t *__restrict__ padj, int32_t nv,
    const unsigned long long *__restrict__ sc, TaskRec *__restrict__ tasks,
    int32_t *__restrict__ bad)
{
    const int32_t hub_lo = (int32_t)sc[1];
    if (hub_lo >= nv) return;
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t e_lo = pcoff[hub_lo] + toff[hub_lo];
    const int64_t n = pcoff[nv] + toff[nv] - e_lo;  // store entries of the hub rows
    const int64_t nchunks = ceil_div(n, kWave);
    const int64_t wstride = (int64_t)gridDim.x * (blockDim.x / kWave);
    for (int64_t c = (int64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; c < nchunks;
         c += wstride) {
        const int64_t i0 = c * kWave, ilast = min(i0 + kWave, n) - 1;
        int32_t wf = 0, wl = 0;
        if (lane == 0) wf = hub_of(pcoff, toff, e_lo + i0, hub_lo, nv - 1);
        if (lane == 1) wl = hub_of(pcoff, toff, e_lo + ilast, hub_lo, nv - 1);
        wf = __shfl(wf, 0, kWave);
        wl = __shfl(wl, 1, kWave);
        const int64_t i = i0 + lane;
        if (i > ilast) continue;
        const int32_t w = wf == wl ? wf : hub_of(pcoff, toff, e_lo + i, wf, wl);
        const int64_t k = e_lo + i - (pcoff[w] + toff[w]);
        const int32_t v = order[w];
        const int64_t rs = ocoff[v];
        const int64_t e = rs + k;  // compact store-order index of the entry
        if ((mask[e >> 6] >> (e & 63)) & 1ull) continue;  // ranked above w: not a task
        const int64_t slot = k - (plus_prefix(cscan, mask, e) - plus_prefix(cscan, mask, rs));
        const int32_t a = rnk[pool[boff[v] + k]];
        const int64_t pa = pcoff[a], pn = pcoff[a + 1] - pa;
        // w is a hub and N+(a) is sorted by rank: w lies among its last
        // nv - hub_lo entries (the hubs), so the search starts there
        int64_t lo = max<int64_t>(0, pn - (nv - hub_lo)), hi = pn;
        while (lo < hi) {
            const int64_t mid = (lo + hi) >> 1;
            if (padj[pa + mid] < w) lo = mid + 1; else hi = mid;
        }
        if (lo < pn && padj[pa + lo] == w)
            tasks[toff[w] + slot] = TaskRec::make(pa + lo + 1, (int32_t)(pn - lo - 1));
        else
            atomicOr(bad, 1);  // rows are not symmetric
    }
}

// First hub rank: degrees ascend with rank, so the ranks of degree at least
// min_deg are [sc[1], nv).  (With graph-major ranks of batched graphs they
// ascend per graph only; the search then lands on some rank, and owners above
// it take the row-built task lists, which are correct for any owner.)  One
// thread.
__global__ void k_hub_lo(const VInfo *__restrict__ vi, int32_t nv, int32_t min_deg,
                         unsigned long long *__restrict__ sc)
{
    int32_t lo = 0, hi = nv;
    while (lo < hi) {
        const int32_t mid = (lo + hi) >> 1;
        if (vi[mid].len < min_deg) lo = mid + 1; else hi = mid;
    }
    sc[1] = (unsigned long long)lo;
    sc[2] = 0ull;
}

// Every N- part below the hubs must have been filled exactly, and the plus
// entries that point to hubs must be as many as the hubs' tasks (each of
// which k_orient_transpose_hubs matched to a distinct plus entry); both hold
// for symmetric rows.
__global__ void k_check_cursor(const int32_t *__restrict__ cursor, const VInfo *__restrict__ vi,
                               const int64_t *__restrict__ toff, int32_t nv,
                               const unsigned long long *__restrict__ sc, int32_t *__restrict__ bad)
{
    const int32_t hub_lo = (int32_t)sc[1];
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < hub_lo;
         r += (int64_t)gridDim.x * blockDim.x)
        if (cursor[r] != vi[r].dm) atomicOr(bad, 1);
    if (blockIdx.x == 0 && threadIdx.x == 0 &&
        sc[2] != (unsigned long long)(toff[nv] - toff[hub_lo]))
        atomicOr(bad, 1);
}

// ---------------------------------------------- partitioned transpose --
//
// The task list of owner w is [toff[w], toff[w+1]) of the task array, so a
// run of consecutive owners owns one contiguous region.  Instead of one
// random 8-byte append (and one cursor atomic) per plus entry, the records
// reach their owners in three passes that each write whole runs:
//  1. coarse: every tile of kTpTile plus entries is split in LDS by coarse
//     bucket (a run of owners); one atomic per (tile, bucket) reserves the
//     bucket's run in its region of a staging array (record and owner), and
//     the tile writes its runs contiguously;
//  2. fine: every tile of a coarse region is split the same way into the fine
//     buckets of that coarse bucket, written into the task array itself (and
//     the owners beside it);
//  3. local: a fine bucket of several owners (fewer than 2 * kTpWin tasks) is
//     counting-sorted by owner in LDS and written back in place.  A fine
//     bucket of one owner is final after pass 2.
// Fine buckets are the owners whose task lists start in one window of kTpWin
// tasks, except that an owner of more than kTpWin tasks is a bucket of its
// own; a coarse bucket is a run of G consecutive fine buckets (G about the
// square root of their number, so that both splits have at most kTpMaxB
// buckets).  The order of an owner's tasks is arbitrary, as with appends.
// Hubs (ranks from sc[1] up) keep the row-built lists of
// k_orient_transpose_hubs.
#ifndef TC_TP_TILE
#define TC_TP_TILE 8192
#endif
#ifndef TC_TP_THREADS
#define TC_TP_THREADS 1024
#endif
#ifndef TC_TP_MAXB
#define TC_TP_MAXB 2048
#endif
constexpr int kTpWin = 4096;
constexpr int kTpTile = TC_TP_TILE;
constexpr int kTpThreads = TC_TP_THREADS;
constexpr int kTpPer = kTpTile / kTpThreads;
constexpr int kTpMaxB = TC_TP_MAXB;
static_assert(kTpMaxB <= 2 * kTpThreads && kTpMaxB <= 2048, "two buckets per thread in the scans");
constexpr int kTpLocal = 2 * kTpWin;
constexpr int64_t kTpDrop = INT64_MIN;

// Fine-bucket starts: owner w < hub_lo starts a bucket when it is the first,
// holds more than kTpWin tasks, follows such an owner, or starts its list in
// another window than the owner before it.
__global__ void k_tp_flags(const int64_t *__restrict__ toff, const unsigned long long *__restrict__ sc,
                           int32_t nv, int32_t *__restrict__ flag)
{
    const int64_t hub_lo = (int64_t)min(sc[1], (unsigned long long)nv);
    for (int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; w < nv;
         w += (int64_t)gridDim.x * blockDim.x) {
        int32_t f = 0;
        if (w < hub_lo) {
            const int64_t t0 = toff[w];
            f = w == 0 || toff[w + 1] - t0 > kTpWin;
            if (!f) {
                const int64_t tb = toff[w - 1];
                f = t0 - tb > kTpWin || t0 / kTpWin != tb / kTpWin;
            }
        }
        flag[w] = f;
    }
}

// fb[i] = first owner of fine bucket i; fb[F] = hub_lo.
__global__ void k_tp_bounds(const int32_t *__restrict__ flag, const int64_t *__restrict__ fpos,
                            const unsigned long long *__restrict__ sc, int32_t nv,
                            int32_t *__restrict__ fb)
{
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (int64_t w = gid; w < nv; w += (int64_t)gridDim.x * blockDim.x)
        if (flag[w]) fb[fpos[w]] = (int32_t)w;
    if (gid == 0) fb[fpos[nv]] = (int32_t)min(sc[1], (unsigned long long)nv);
}

// One block of 1024 threads: ctl = {F, G, K (coarse buckets), pass-2 tiles},
// tiles[k] = first pass-2 tile of coarse bucket k (tiles[K] = their total),
// and the coarse cursors zeroed.
__global__ __launch_bounds__(1024) void k_tp_plan(const int64_t *__restrict__ fpos, int32_t nv,
                                                  const int32_t *__restrict__ fb,
                                                  const int64_t *__restrict__ toff,
                                                  int64_t *__restrict__ ctl, int32_t *__restrict__ tiles,
                                                  unsigned long long *__restrict__ ccur)
{
    __shared__ int32_t wsum[1024 / kWave];
    const int t = threadIdx.x, lane = t & (kWave - 1), wid = t / kWave;
    const int64_t F = fpos[nv];
    int64_t G = 1;
    if (F > 0) {
        G = max<int64_t>(1, (int64_t)ceil(sqrt((double)F)));
        if (ceil_div(F, G) > kTpMaxB) G = ceil_div(F, kTpMaxB);
    }
    const int64_t K = F > 0 ? ceil_div(F, G) : 0;
    int32_t c[2];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const int64_t k = 2 * t + q;
        c[q] = 0;
        if (k < K) {
            const int64_t size = toff[fb[min(G * (k + 1), F)]] - toff[fb[G * k]];
            c[q] = (int32_t)ceil_div(size, kTpTile);
        }
        if (k < kTpMaxB) ccur[k] = 0ull;
    }
    const int32_t sum = c[0] + c[1];
    const int32_t incl = wave_incl_scan_i32(sum);
    if (lane == kWave - 1) wsum[wid] = incl;
    __syncthreads();
    int32_t before = 0, total = 0;
    for (int k = 0; k < 1024 / kWave; ++k) {
        before += k < wid ? wsum[k] : 0;
        total += wsum[k];
    }
    const int32_t e0 = before + incl - sum;
    if (2 * t < K) tiles[2 * t] = e0;
    if (2 * t + 1 < K) tiles[2 * t + 1] = e0 + c[0];
    if (t == 0) {
        tiles[K] = total;
        ctl[0] = F;
        ctl[1] = G;
        ctl[2] = K;
        ctl[3] = total;
    }
}

// tmap[i] = the coarse bucket of pass-2 tile i (tiles[k] <= i < tiles[k+1]);
// one block per coarse bucket.
__global__ void k_tp_tilemap(const int64_t *__restrict__ ctl, const int32_t *__restrict__ tiles,
                             int32_t *__restrict__ tmap)
{
    const int64_t k = blockIdx.x;
    if (k >= ctl[2]) return;
    for (int32_t i = tiles[k] + threadIdx.x; i < tiles[k + 1]; i += blockDim.x) tmap[i] = (int32_t)k;
}

struct TpShared {
    int32_t bnd[kTpMaxB + 1];  // owner bounds of the tile's buckets
    int32_t cnt[kTpMaxB];      // records per bucket, then their local offsets
    int64_t dst[kTpMaxB];      // where the bucket's run goes, less its local offset
    alignas(16) uint64_t rec[kTpTile];
    alignas(16) int32_t own[kTpTile];  // (read as int4 by k_tp_coarse)
    uint16_t bk[kTpTile];
    int32_t wsum[kTpThreads / kWave];
};

// For every record k with bk[k] >= 0: the largest b in [0, nb) with
// bnd[b] <= own[k] (0 when none), into bk[k].  The searches of a thread's
// records advance in lockstep, so their LDS reads are in flight together.
__device__ inline void tp_buckets(const int32_t *bnd, int32_t nb, const int32_t (&own)[kTpPer],
                                  int32_t (&bk)[kTpPer])
{
    int32_t pos[kTpPer];
#pragma unroll
    for (int k = 0; k < kTpPer; ++k) pos[k] = 0;
    int32_t step = 1;
    while (2 * step < nb) step *= 2;
    for (; step > 0; step >>= 1) {
#pragma unroll
        for (int k = 0; k < kTpPer; ++k) {
            const int32_t q = pos[k] + step;
            if (q < nb && bnd[q] <= own[k]) pos[k] = q;
        }
    }
#pragma unroll
    for (int k = 0; k < kTpPer; ++k)
        if (bk[k] >= 0) bk[k] = pos[k];
}

// The split of one tile: every thread holds up to kTpPer records with their
// bucket (bk < 0: none) and rank in it (from atomics on S.cnt, complete on
// entry up to the barrier below).  Reserves each bucket's run with one
// atomic on cur[b], whose region is [toff[bnd[b]], toff[bnd[b+1]]); a run
// past its region is dropped and flags the rows as not symmetric.
__device__ inline void tp_split(TpShared &S, int32_t nb, const uint64_t (&rec)[kTpPer],
                                const int32_t (&own)[kTpPer], const int32_t (&bk)[kTpPer],
                                const int32_t (&rank)[kTpPer], const int64_t *__restrict__ toff,
                                unsigned long long *__restrict__ cur, uint64_t *__restrict__ rec_out,
                                int32_t *__restrict__ own_out, int32_t *__restrict__ bad)
{
    const int t = threadIdx.x, lane = t & (kWave - 1), wid = t / kWave;
    __syncthreads();  // the counts are complete
    int32_t c[2];
#pragma unroll
    for (int q = 0; q < 2; ++q) c[q] = 2 * t + q < nb ? S.cnt[2 * t + q] : 0;
    const int32_t sum = c[0] + c[1];
    const int32_t incl = wave_incl_scan_i32(sum);
    if (lane == kWave - 1) S.wsum[wid] = incl;
    __syncthreads();
    int32_t before = 0, total = 0;
    for (int k = 0; k < kTpThreads / kWave; ++k) {
        before += k < wid ? S.wsum[k] : 0;
        total += S.wsum[k];
    }
    int32_t off = before + incl - sum;
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const int32_t b = 2 * t + q;
        if (b < nb) {
            if (c[q] > 0) {
                const int64_t start = toff[S.bnd[b]], size = toff[S.bnd[b + 1]] - start;
                const unsigned long long r = atomicAdd(&cur[b], (unsigned long long)c[q]);
                if ((int64_t)r + c[q] <= size) {
                    S.dst[b] = start + (int64_t)r - off;
                } else {
                    S.dst[b] = kTpDrop;
                    atomicOr(bad, 1);  // more records than the owners' tasks
                }
            }
            S.cnt[b] = off;
        }
        off += c[q];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kTpPer; ++k) {
        if (bk[k] < 0) continue;
        const int32_t p = S.cnt[bk[k]] + rank[k];
        S.rec[p] = rec[k];
        S.own[p] = own[k];
        S.bk[p] = (uint16_t)bk[k];
    }
    __syncthreads();
    for (int32_t p = t; p < total; p += kTpThreads) {
        const int64_t d = S.dst[S.bk[p]];
        if (d != kTpDrop) {
            rec_out[d + p] = S.rec[p];
            own_out[d + p] = S.own[p];
        }
    }
}

// Pass 1: the plus entries of one tile into their coarse buckets' regions of
// the staging arrays (task record and owner).  Entries that point to hubs
// are counted in sc[2] instead.  An entry's suffix length needs the end of
// its row: the rows that end inside the tile mark their last entry (in the
// owner staging area, free until the split), a suffix minimum spreads each
// mark over its row, and the entries of a row that runs past the tile take
// that row's end.
__global__ __launch_bounds__(kTpThreads) void k_tp_coarse(
    const int64_t *__restrict__ pcoff, const int32_t *__restrict__ pcrow,
    const int32_t *__restrict__ padj, const int64_t *__restrict__ toff,
    const int32_t *__restrict__ fb, const int64_t *__restrict__ ctl, int32_t nv, int64_t nplus,
    unsigned long long *__restrict__ sc, unsigned long long *__restrict__ ccur,
    uint64_t *__restrict__ rec_out, int32_t *__restrict__ own_out, int32_t *__restrict__ bad)
{
    __shared__ TpShared S;
    __shared__ int32_t end_last;
    const int t = threadIdx.x, lane = t & (kWave - 1), wid = t / kWave;
    const int64_t F = ctl[0], G = ctl[1];
    const int32_t K = (int32_t)ctl[2];
    const int32_t hub_lo = (int32_t)min(sc[1], (unsigned long long)nv);
    const int64_t E0 = (int64_t)blockIdx.x * kTpTile, E1 = min(E0 + kTpTile, nplus);
    int32_t own[kTpPer];  // the tile's entries, loaded first: in flight while the rows are marked
#pragma unroll
    for (int k = 0; k < kTpPer; ++k) {
        const int64_t e = E0 + k * kTpThreads + t;
        own[k] = e < E1 ? padj[e] : INT32_MAX;
    }
    int32_t *endm = S.own;  // row ends by tile entry (relative to E0)
    for (int32_t b = t; b <= K; b += kTpThreads) S.bnd[b] = fb[min(G * b, F)];
    for (int32_t b = t; b < K; b += kTpThreads) S.cnt[b] = 0;
    for (int i = t; i < kTpTile; i += kTpThreads) endm[i] = INT32_MAX;
    if (t == 0) end_last = INT32_MAX;
    __syncthreads();
    // rows that overlap the tile: from the row of entry E0 to the row of E1
    const int32_t rlo = pcrow[E0 / kWave];
    const int32_t rhi = E1 < nplus ? pcrow[E1 / kWave] : nv - 1;
    for (int64_t r = rlo + t; r <= rhi; r += kTpThreads) {
        const int64_t b = pcoff[r], e = pcoff[r + 1];
        if (e > E0 && e <= E1) endm[e - 1 - E0] = (int32_t)(e - E0);
        else if (e > E1 && b < E1) end_last = (int32_t)(e - E0);  // runs past the tile
    }
    __syncthreads();
    {  // suffix minimum: thread t owns entries [8t, 8t + 8)
        static_assert(kTpPer == 8, "two int4 per thread");
        int4 *e4 = reinterpret_cast<int4 *>(endm);
        const int4 a = e4[2 * t], c = e4[2 * t + 1];
        int32_t v[8] = {a.x, a.y, a.z, a.w, c.x, c.y, c.z, c.w};
#pragma unroll
        for (int k = 6; k >= 0; --k) v[k] = min(v[k], v[k + 1]);
        int32_t x = v[0];  // minimum over the lanes from this one up
#pragma unroll
        for (int d = 1; d < kWave; d <<= 1) {
            const int32_t y = __shfl_down(x, d, kWave);
            if (lane + d < kWave) x = min(x, y);
        }
        if (lane == 0) S.wsum[wid] = x;
        __syncthreads();
        int32_t after = __shfl_down(x, 1, kWave);
        if (lane == kWave - 1) after = INT32_MAX;
        for (int q = wid + 1; q < kTpThreads / kWave; ++q) after = min(after, S.wsum[q]);
#pragma unroll
        for (int k = 0; k < 8; ++k) v[k] = min(v[k], after);
        e4[2 * t] = make_int4(v[0], v[1], v[2], v[3]);
        e4[2 * t + 1] = make_int4(v[4], v[5], v[6], v[7]);
        __syncthreads();
    }
    uint64_t rec[kTpPer];
    int32_t bk[kTpPer], rank[kTpPer];
    unsigned long long nhub = 0;
#pragma unroll
    for (int k = 0; k < kTpPer; ++k) {
        const int32_t i = k * kTpThreads + t;
        const int64_t e = E0 + i;
        bk[k] = -1;
        rec[k] = 0;
        rank[k] = 0;
        if (e >= E1) continue;
        if (own[k] >= hub_lo) {
            ++nhub;  // filled by k_orient_transpose_hubs
            continue;
        }
        bk[k] = 0;
        int32_t end = endm[i];
        if (end == INT32_MAX) end = end_last;
        rec[k] = TaskRec::make(e + 1, end - i - 1).v;
    }
    tp_buckets(S.bnd, K, own, bk);
#pragma unroll
    for (int k = 0; k < kTpPer; ++k)
        if (bk[k] >= 0) rank[k] = atomicAdd(&S.cnt[bk[k]], 1);
    tp_split(S, K, rec, own, bk, rank, toff, ccur, rec_out, own_out, bad);
    const unsigned long long part[1] = {nhub};
    block_add_u64<1>(part, sc + 2);
}

// Pass 2: one tile of a coarse region into the fine buckets of that coarse
// bucket (records into the task array, owners beside them).
__global__ __launch_bounds__(kTpThreads) void k_tp_fine(
    const uint64_t *__restrict__ rec_in, const int32_t *__restrict__ own_in,
    const int64_t *__restrict__ toff, const int32_t *__restrict__ fb,
    const int64_t *__restrict__ ctl, const int32_t *__restrict__ tiles,
    const int32_t *__restrict__ tmap, unsigned long long *__restrict__ fcur,
    uint64_t *__restrict__ rec_out, int32_t *__restrict__ own_out, int32_t *__restrict__ bad)
{
    __shared__ TpShared S;
    const int t = threadIdx.x;
    const int64_t F = ctl[0], G = ctl[1], ntiles = ctl[3];
    const int64_t bid = blockIdx.x;
    if (bid >= ntiles) return;
    const int64_t k = tmap[bid];  // the coarse bucket
    const int64_t f0 = G * k;
    const int32_t nf = (int32_t)min(G, F - f0);
    const int64_t r1 = toff[fb[f0 + nf]];
    const int64_t p0 = toff[fb[f0]] + (bid - tiles[k]) * kTpTile;
    const int64_t p1 = min(p0 + kTpTile, r1);
    uint64_t rec[kTpPer];
    int32_t own[kTpPer], bk[kTpPer], rank[kTpPer];
#pragma unroll
    for (int q = 0; q < kTpPer; ++q) {  // in flight while the bounds are set up
        const int64_t p = p0 + (int64_t)q * kTpThreads + t;
        bk[q] = p < p1 ? 0 : -1;
        rec[q] = p < p1 ? rec_in[p] : 0;
        own[q] = p < p1 ? own_in[p] : 0;
        rank[q] = 0;
    }
    for (int32_t b = t; b <= nf; b += kTpThreads) S.bnd[b] = fb[f0 + b];
    for (int32_t b = t; b < nf; b += kTpThreads) S.cnt[b] = 0;
    __syncthreads();
    tp_buckets(S.bnd, nf, own, bk);
#pragma unroll
    for (int q = 0; q < kTpPer; ++q)
        if (bk[q] >= 0) rank[q] = atomicAdd(&S.cnt[bk[q]], 1);
    tp_split(S, nf, rec, own, bk, rank, toff, fcur + f0, rec_out, own_out, bad);
}

// Pass 3: every fine bucket of several owners, counting-sorted by owner in
// LDS: the record of owner w with rank s among w's records goes to
// toff[w] + s.  Owner w's key toff[w] - (bucket start) is below kTpWin (all
// owners of the bucket start their lists in one window), so counters by key
// need no owner index.  A record of an owner outside the bucket, of an owner
// without tasks, or past its owner's task count flags the rows.
__global__ __launch_bounds__(kTpThreads) void k_tp_local(const int64_t *__restrict__ toff,
                                                         const int32_t *__restrict__ fb,
                                                         const int64_t *__restrict__ ctl,
                                                         uint64_t *__restrict__ tasks,
                                                         const int32_t *__restrict__ own,
                                                         int32_t *__restrict__ bad)
{
    __shared__ int32_t lc[kTpWin];
    __shared__ uint64_t st[kTpLocal];
    const int t = threadIdx.x;
    const int64_t F = ctl[0];
    for (int64_t f = blockIdx.x; f < F; f += gridDim.x) {
        const int32_t w0 = fb[f], w1 = fb[f + 1];
        if (w1 - w0 <= 1) continue;  // one owner: its records are final
        const int64_t rs = toff[w0], n = toff[w1] - rs;
        if (n > kTpLocal) {  // not possible for buckets built by k_tp_flags
            if (t == 0) atomicOr(bad, 1);
            continue;
        }
        for (int i = t; i < kTpWin; i += kTpThreads) lc[i] = 0;
        __syncthreads();
        bool ok = true;
        for (int64_t i = t; i < n; i += kTpThreads) {
            const uint64_t r = tasks[rs + i];
            const int32_t w = own[rs + i];
            bool good = w >= w0 && w < w1;
            if (good) {
                const int64_t a = toff[w], d = toff[w + 1] - a, key = a - rs;
                good = d > 0 && key >= 0 && key < kTpWin;
                if (good) {
                    const int32_t s = atomicAdd(&lc[key], 1);
                    good = s < d;
                    if (good) st[key + s] = r;
                }
            }
            ok = ok && good;
        }
        if (!ok) atomicOr(bad, 1);
        __syncthreads();
        for (int64_t i = t; i < n; i += kTpThreads) tasks[rs + i] = st[i];
        __syncthreads();
    }
}

// Every coarse and fine region received exactly its owners' task count
// (with the per-record checks of pass 3 this is the per-owner count check),
// and the plus entries that point to hubs are as many as the hubs' tasks.
__global__ void k_tp_check(const int64_t *__restrict__ toff, const int32_t *__restrict__ fb,
                           const int64_t *__restrict__ ctl, const unsigned long long *__restrict__ ccur,
                           const unsigned long long *__restrict__ fcur, int32_t nv,
                           const unsigned long long *__restrict__ sc, int32_t *__restrict__ bad)
{
    const int64_t F = ctl[0], G = ctl[1], K = ctl[2];
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = gid; i < F; i += stride)
        if ((int64_t)fcur[i] != toff[fb[i + 1]] - toff[fb[i]]) atomicOr(bad, 1);
    for (int64_t k = gid; k < K; k += stride)
        if ((int64_t)ccur[k] != toff[fb[min(G * (k + 1), F)]] - toff[fb[G * k]]) atomicOr(bad, 1);
    if (gid == 0) {
        const int64_t hub_lo = (int64_t)min(sc[1], (unsigned long long)nv);
        if (sc[2] != (unsigned long long)(toff[nv] - toff[hub_lo])) atomicOr(bad, 1);
    }
}

__global__ void k_direct_vinfo(const int64_t *__restrict__ boff, const int32_t *__restrict__ len,
                               int32_t nv, VInfo *__restrict__ vi)
{
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < nv;
         v += (int64_t)gridDim.x * blockDim.x) {
        VInfo x;
        x.rb = boff[v];
        x.len = len[v];
        x.dm = 0;
        vi[v] = x;
    }
}

// ------------------------------------------------------------- planning --

struct PlanParams {
    int32_t lmax[3];     // owner-list capacity of each bin's kernel
    int32_t nw[3];       // wavefronts per block of each bin's kernel
    int32_t bin_max[2];  // |owner| <= bin_max[0] -> bin 0, <= bin_max[1] -> bin 1, else 2
    int32_t job_items;   // target estimated items per job (per wavefront)
    int32_t flush_items; // minimum items per owner slot in a job
};

__device__ inline int32_t est_items(int32_t L, int32_t pl)
{
    // cheapest of streaming the task list (pl) and searching the owner list (L*log pl)
    const int32_t lg = max(1, 32 - __clz(pl));  // smallest lg >= 1 with 2^lg > pl
    int64_t bs = (int64_t)L * lg;
    return (int32_t)min<int64_t>(pl, bs) + 4;
}

__device__ inline int owner_bin(int32_t L, const PlanParams &P)
{
    return L <= P.bin_max[0] ? 0 : (L <= P.bin_max[1] ? 1 : 2);
}

// Per-owner estimated work (sum over its tasks); per-bin stats.
template <bool ORIENTED>
// Entries are the task array (oriented) or the store rows (direct), with
// their row offsets coff and chunk rows crow.
__global__ void k_owner_work(const int64_t *__restrict__ coff, const int32_t *__restrict__ crow,
                             const VInfo *__restrict__ vi,
                             const int32_t *__restrict__ adj, const TaskRec *__restrict__ tasks,
                             int32_t nv, int64_t nnz, PlanParams P,
                             unsigned long long *__restrict__ work,
                             unsigned long long *__restrict__ bstats)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t nchunks = ceil_div(nnz, kWave);
    const int64_t wstride = (int64_t)gridDim.x * (blockDim.x / kWave);
    // per-lane bin statistics, reduced once at the end
    unsigned long long st_t[3] = {0, 0, 0}, st_b[3] = {0, 0, 0}, st_i[3] = {0, 0, 0};
    // one chunk: c, its row b per lane, the task suffix length tl (oriented)
    auto chunk = [&](int64_t c, int32_t b, int32_t tl) {
        const int64_t e = c * kWave + lane;
        uint32_t w = 0;
        if (e < nnz) {
            VInfo vb = vi[b];
            int32_t L = ORIENTED ? vb.len - vb.dm : vb.len;
            if (L > 0) {
                int32_t pl;
                bool ok;
                if (ORIENTED) {
                    // only the suffix of N+(a) ranked above the owner can hit
                    pl = tl;
                    ok = pl > 0;
                } else {
                    const int32_t a = adj[vb.rb + (e - coff[b])];
                    const VInfo va = vi[a];
                    pl = va.len;
                    ok = a != b && va.len >= 2 && vb.len >= 2;
                }
                if (ok) {
                    w = (uint32_t)est_items(min(L, P.lmax[2]), pl);
                    const int bin = owner_bin(L, P);
#pragma unroll
                    for (int q = 0; q < 3; ++q)
                        if (bin == q) {
                            st_t[q] += 1;
                            st_b[q] += 4ull * (unsigned long long)(L + pl);
                            st_i[q] += (unsigned long long)pl;
                        }
                }
            }
        }
        // segmented (by owner row) inclusive scan of w across the wave; a
        // chunk's sum fits 32 bits (each item estimate is below 2^25)
        uint32_t x = w;
#pragma unroll
        for (int d = 1; d < kWave; d <<= 1) {
            uint32_t y = __shfl_up(x, d, kWave);
            int32_t rb = __shfl_up(b, d, kWave);
            if (lane >= d && rb == b) x += y;
        }
        int32_t bnext = __shfl_down(b, 1, kWave);
        bool tail = (lane == kWave - 1) || (bnext != b) || (e + 1 >= nnz);
        if (e < nnz && tail && x) atomicAdd(&work[b], (unsigned long long)x);
    };
    // two chunks per pass (c and c + wstride): the task records of both and
    // their row searches in flight together
    for (int64_t c = (int64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; c < nchunks;
         c += 2 * wstride) {
        const int64_t c2 = c + wstride;
        const bool two = c2 < nchunks;  // (wave-uniform)
        const int64_t e = c * kWave + lane, e2 = c2 * kWave + lane;
        const int32_t tl = (ORIENTED && e < nnz) ? tasks[e].len() : 0;
        const int32_t tl2 = (ORIENTED && two && e2 < nnz) ? tasks[e2].len() : 0;
        int32_t b, b2;
        wave_rows_of2(coff, crow, nv, c, two ? c2 : c, nnz, b, b2);
        chunk(c, b, tl);
        if (two) chunk(c2, b2, tl2);
    }
    if (bstats) {
        // bstats layout: tasks/bytes pairs per bin, then items per bin
        const unsigned long long v[9] = {st_t[0], st_b[0], st_t[1], st_b[1], st_t[2],
                                         st_b[2], st_i[0], st_i[1], st_i[2]};
        block_add_u64<9>(v, bstats);
    }
}

// Jobs per owner and bin; jstats[b] (optional) gets the owner-tile entries
// the jobs of bin b load.
template <bool ORIENTED>
__global__ void k_plan_jobs(const VInfo *__restrict__ vi, const unsigned long long *__restrict__ work,
                            int32_t nv, PlanParams P, int32_t *__restrict__ njobs /* [3][nv] */,
                            unsigned long long *__restrict__ jstats)
{
    const int lane = threadIdx.x & (kWave - 1);
    unsigned long long tile_items[3] = {0, 0, 0};
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t v0 = (int64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~(kWave - 1)); v0 < nv;
         v0 += stride) {
        const int64_t v = v0 + lane;
        if (v >= nv) continue;
        VInfo x = vi[v];
        int32_t L = ORIENTED ? x.len - x.dm : x.len;
        int32_t T = ORIENTED ? x.dm : x.len;
        int32_t nj[3] = {0, 0, 0};
        unsigned long long w = work[v];
        if (L > 0 && T > 0 && w > 0 && (ORIENTED || x.len >= 2)) {
            int bin = owner_bin(L, P);
            int64_t tiles = ceil_div(L, P.lmax[bin]);
            int64_t per_tile = (int64_t)(w / (unsigned long long)tiles) + 1;
            // a job amortises building its LDS table over at least 16 items per slot
            // Each job ends by flushing up to one global atomic per owner slot,
            // so a job gets at least flush_items items per slot.
            int64_t grain = max<int64_t>((int64_t)P.job_items * P.nw[bin],
                                         (int64_t)P.flush_items * min(L, P.lmax[bin]));
            int64_t split = ceil_div(per_tile, grain);
            if (split > T) split = T;
            if (split < 1) split = 1;
            nj[bin] = (int32_t)(split * tiles);
#pragma unroll
            for (int q = 0; q < 3; ++q)
                if (q == bin) tile_items[q] += (unsigned long long)split * (unsigned long long)L;
        }
        njobs[v] = nj[0];
        njobs[(int64_t)nv + v] = nj[1];
        njobs[2 * (int64_t)nv + v] = nj[2];
    }
    if (jstats) block_add_u64<3>(tile_items, jstats);
}

template <bool ORIENTED>
__global__ void k_emit_jobs(const VInfo *__restrict__ vi, const int32_t *__restrict__ njobs,
                            const int64_t *__restrict__ jscan /* [3][nv+1] */, int32_t nv,
                            PlanParams P, Job *__restrict__ jobs0, Job *__restrict__ jobs1,
                            Job *__restrict__ jobs2)
{
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < nv;
         v += (int64_t)gridDim.x * blockDim.x) {
        VInfo x = vi[v];
        int32_t L = ORIENTED ? x.len - x.dm : x.len;
        int32_t T = ORIENTED ? x.dm : x.len;
        for (int bin = 0; bin < 3; ++bin) {
            int32_t nj = njobs[(int64_t)bin * nv + v];
            if (!nj) continue;
            Job *out = bin == 0 ? jobs0 : (bin == 1 ? jobs1 : jobs2);
            int64_t base = jscan[(int64_t)bin * (nv + 1) + v];
            int32_t tiles = (int32_t)ceil_div(L, P.lmax[bin]);
            int32_t split = nj / tiles;
            int64_t q = 0;
            for (int32_t t = 0; t < tiles; ++t)
                for (int32_t s = 0; s < split; ++s) {
                    Job j;
                    j.owner = (int32_t)v;
                    j.t_begin = (int32_t)(((int64_t)T * s) / split);
                    j.t_end = (int32_t)(((int64_t)T * (s + 1)) / split);
                    j.tile = t;
                    out[base + q++] = j;
                }
        }
    }
}

// outr[r] += 2 * (sum of pcnt over N+(r)): the task-vertex share of the
// oriented count, one segmented wave sum per 64 entries of the plus array.
__global__ void k_plus_row_sum(const int64_t *__restrict__ pcoff, const int32_t *__restrict__ pcrow,
                               const uint32_t *__restrict__ pcnt, int32_t nv, int64_t nplus,
                               uint32_t *__restrict__ outr)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t nchunks = ceil_div(nplus, kWave);
    const int64_t wstride = (int64_t)gridDim.x * (blockDim.x / kWave);
    auto chunk = [&](int64_t e, int32_t r, uint32_t x) {
#pragma unroll
        for (int d = 1; d < kWave; d <<= 1) {
            const uint32_t y = __shfl_up(x, d, kWave);
            const int32_t ry = __shfl_up(r, d, kWave);
            if (lane >= d && ry == r) x += y;
        }
        const int32_t rn = __shfl_down(r, 1, kWave);
        const bool tail = lane == kWave - 1 || rn != r || e + 1 >= nplus;
        if (e < nplus && tail && x) atomicAdd(&outr[r], 2u * x);
    };
    // two chunks per pass (c and c + wstride), their counts and row searches
    // in flight together
    for (int64_t c = (int64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; c < nchunks;
         c += 2 * wstride) {
        const int64_t c2 = c + wstride;
        const bool two = c2 < nchunks;  // (wave-uniform)
        const int64_t e = c * kWave + lane, e2 = c2 * kWave + lane;
        const uint32_t x = e < nplus ? pcnt[e] : 0u;
        const uint32_t x2 = (two && e2 < nplus) ? pcnt[e2] : 0u;
        int32_t r, r2;
        wave_rows_of2(pcoff, pcrow, nv, c, two ? c2 : c, nplus, r, r2);
        chunk(e, r, x);
        if (two) chunk(e2, r2, x2);
    }
}

// out[order[r]] += outr[r]: the counts credited by rank added to the
// per-vertex counts (every vertex once, after the count kernels).
__global__ void k_rank_to_vertex(const int32_t *__restrict__ order, int32_t nv,
                                 const uint32_t *__restrict__ outr, uint32_t *__restrict__ out)
{
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < nv;
         r += (int64_t)gridDim.x * blockDim.x)
        out[order[r]] += outr[r];
}

// gid[r] = graph holding vertex order[r] (or r): last g with voff[g] <= vertex.
__global__ void k_graph_of(const int64_t *__restrict__ voff, int32_t G,
                           const int32_t *__restrict__ order, int32_t nv, int32_t *__restrict__ gid)
{
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < nv;
         r += (int64_t)gridDim.x * blockDim.x) {
        const int64_t v = order ? order[r] : r;
        int32_t lo = 0, hi = G - 1;
        while (lo < hi) {
            const int32_t mid = (lo + hi + 1) >> 1;
            if (voff[mid] <= v) lo = mid; else hi = mid - 1;
        }
        gid[r] = lo;
    }
}

// ---------------------------------------------------------- count kernel --

struct CountArgs {
    const int32_t *adj;
    const VInfo *vi;
    const TaskRec *tasks;  // oriented: task vertex and suffix of each task
    const int64_t *toff;   // oriented: first task of each owner
    const int32_t *order;  // oriented: vertex of each rank
    uint32_t *pcnt;        // oriented: per N+ entry, the hits of its task
    uint32_t *out;         // per-vertex counts
    // oriented: per-vertex counts by rank, credited by k_count_long and
    // k_plus_row_sum without an order[] lookup and added to out at the end
    // (k_count keeps crediting out: its hot counters, adjacent in rank order,
    // would share cache lines there -- mid bin 20 -> 29 ms at R-MAT 24)
    uint32_t *outr;
    unsigned long long *global;
    unsigned long long *strat;  // [3] tasks per strategy (may be null)
    unsigned long long *strat_mp;  // tasks run as a merge path (may be null)
    const uint8_t *table;       // learned strategy table (policy LEARNED)
    int32_t table_dim;
    int32_t policy;
    int32_t ratio_x16;
    int32_t mpath_min;   // fixed heuristic: balanced pairs of at least this length merge-path
    int32_t skip_lower;  // CSR convention: entries with dest < src are skipped
    // batched graphs: graph g owns vertices [gvoff[g], gvoff[g+1]); a job's
    // hits are added to per_graph[g] of its owner (triangles never span graphs)
    const int32_t *gid;  // graph of each owner (rank or vertex id, as jobs name them)
    unsigned long long *per_graph;
    // largest bin, oriented: hash tasks are left to k_count_long, which adds
    // its tasks, streamed entries and owner-tile entries to long_stats[0..2]
    // (bitmap tiles) or [3..5] (hash-table tiles)
    int32_t long_external;
    unsigned long long *long_stats;
    // largest bin, oriented (every kernel of it): the fixed and learned
    // policies stream every task through the tile -- a task of another
    // strategy makes k_count rebuild its job's tile, a per-job cost the
    // per-task timings behind the table never see (forced policies still hold)
    int32_t hash_only;
    // zeroed job counters of the largest bin's launches: [0] the bitmap
    // launch's next job, [1] the length of the list hjobs of the jobs it left
    // to the hash-table launch, [2] that launch's next entry of the list
    unsigned long long *job_ctr;
    int64_t *hjobs;  // room for every job of the largest bin (or null)
};

__device__ inline int lg_bucket(int32_t x)
{
    // floor(log2(x)) + 1 for x >= 1, 0 for x == 0
    return x <= 0 ? 0 : 32 - __clz(x);
}

__device__ inline int choose_strategy(int32_t L, int32_t pl, const CountArgs &A)
{
    switch (A.policy) {
    case TC_POLICY_MERGE: return S_MERGE;
    case TC_POLICY_BSEARCH: return S_BSEARCH;
    case TC_POLICY_HASH: return S_HASH;
    case TC_POLICY_MERGEPATH: return S_MPATH;
    case TC_POLICY_LEARNED:
        if (A.hash_only) return S_HASH;
        if (A.table) {
            int i = min(lg_bucket(L), A.table_dim - 1);
            int j = min(lg_bucket(pl), A.table_dim - 1);
            return A.table[i * A.table_dim + j];
        }
        break;
    default: break;
    }
    if (A.hash_only) return S_HASH;
    // fixed degree-ratio heuristic; merge only when ratio_x16 asks for it:
    // measured on 4096 batched R-MAT graphs, merging the short similar lists
    // it used to take (long <= 4 * short, long <= 96) cost 1.49x against
    // streaming them through the table (lane-sequential walks in lockstep)
    int32_t sh = min(L, pl), lo = max(L, pl);
    if (A.ratio_x16 > 0 && (int64_t)lo * 16 <= (int64_t)sh * A.ratio_x16 && lo <= 96) return S_MERGE;
    if (A.mpath_min > 0 && sh >= A.mpath_min && lo <= 2 * sh) return S_MPATH;
    // searching the owner list into a much longer task list beats streaming it
    if ((int64_t)L * lg_bucket(pl) * 4 < (int64_t)pl) return S_BSEARCH;
    return S_HASH;
}

// buffer resource word 3 for raw 32-bit loads on gfx9-family GPUs (stride 0,
// range checked against num_records in bytes; out-of-range lanes read 0)
constexpr int kBufFlags = 0x00020000;

// Streams the task lists of one wave's 64 tasks (lane t's list: pl entries at
// adj[pb ..]) as one sequence of 64-entry chunks: chunk c belongs to the first
// lane whose inclusive chunk count exceeds c, so chunk owners are wave-uniform
// and a batch of U chunks spans several short tasks.  Loads go through a buffer
// resource over the owning task's remaining entries: lanes past its end (and
// every chunk past the last) read 0 from the range check, with no memory
// access, no branch and no exec mask.  Two batches of registers are used in
// turn, so the next batch's loads are in flight while a batch is probed.  The
// owning task is tracked in scalar registers and looked up again (one ballot,
// a few readlanes) only when a chunk passes its end.  probe(w, hc) probes a
// batch and sets hc[u] to the hits of chunk u over the wave.  ZERO_SAFE: 0 is
// never in the probed set (the oriented layout, whose lists hold ranks above
// their owner); otherwise lanes past a task's end read -1.  Returns the hits of
// this lane's task.  Called by the whole wave.
template <int U, bool ZERO_SAFE, class Probe>
__device__ inline uint32_t stream_tasks(const int32_t *adj, int32_t pl, int64_t pb, Probe &&probe)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int32_t myc = (pl + kWave - 1) / kWave;
    const int32_t cincl = wave_incl_scan_i32(myc);
    const int32_t ctot = __builtin_amdgcn_readlane(cincl, kWave - 1);
    if (ctot == 0) return 0;
    const int32_t lane4 = lane * 4;
    // the loading side's task: the end of its chunk range (INT32_MAX once the
    // last task is reached), the byte address of its next chunk and the bytes
    // left from there.  Chunks are visited in order, so a task is entered at
    // its first chunk.
    int32_t lend = 0, lrem = 0;
    uint64_t lp = 0;
    auto load = [&](int32_t c0, int32_t (&w)[U]) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int32_t c = c0 + u;
            if (c >= lend) {
                const int j = __ffsll((unsigned long long)__ballot(cincl > c)) - 1;
                lend = __builtin_amdgcn_readlane(cincl, j);
                if (lend == ctot) lend = INT32_MAX;
                lrem = __builtin_amdgcn_readlane(pl, j) * 4;
                lp = reinterpret_cast<uint64_t>(adj + readlane_i64(pb, j));
            }
            const __amdgpu_buffer_rsrc_t rs =
                __builtin_amdgcn_make_buffer_rsrc(reinterpret_cast<void *>(lp), 0, max(lrem, 0), kBufFlags);
            int32_t v = __builtin_amdgcn_raw_buffer_load_b32(rs, lane4, 0, 0);
            if (!ZERO_SAFE && lane4 >= lrem) v = -1;
            w[u] = v;
            lp += kWave * 4;
            lrem -= kWave * 4;
        }
    };
    // the probing side's task and its hits so far; a task's chunks are
    // consecutive, so its total is final when the chunk owner changes
    uint32_t mine_h = 0, pacc = 0;
    int32_t pj = 0, pend = 0;
    auto book = [&](int32_t c0, const uint32_t (&hc)[U]) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int32_t c = c0 + u;
            if (c >= pend) {  // INT32_MAX once the last task is reached
                if (pacc && lane == pj) mine_h = pacc;
                pj = __ffsll((unsigned long long)__ballot(cincl > c)) - 1;
                pend = __builtin_amdgcn_readlane(cincl, pj);
                if (pend == ctot) pend = INT32_MAX;
                pacc = 0;
            }
            pacc += hc[u];
        }
    };
    int32_t ba[U], bb[U];
    uint32_t hc[U];
    load(0, ba);
    for (int32_t c0 = 0;; c0 += 2 * U) {
        load(c0 + U, bb);  // all zeros past the last chunk
        probe(ba, hc);
        book(c0, hc);
        if (c0 + U >= ctot) break;
        load(c0 + 2 * U, ba);
        probe(bb, hc);
        book(c0 + U, hc);
        if (c0 + 2 * U >= ctot) break;
    }
    if (pacc && lane == pj) mine_h = pacc;
    return mine_h;
}

// |X cap Y| of two strictly increasing lists by the whole wavefront, as a
// merge path over 64-entry windows (the cooperative count of
// singleIntersection / intersectPath, src/static_triangle_counting/cct.cu:
// 145-199, with the diagonal search of bSearchIP, :91-132).  Each step loads
// the next 64 entries of both lists (coalesced) and every lane finds the
// co-rank of its x in the Y window -- the merge-path split at x's diagonal --
// by a six-step binary search over the window's lanes.  Only values up to
// m = min(last x, last y) are settled by the two windows: a pair x == y <= m
// has both ends in them, so it is counted there and nowhere else (no lane
// hands a match at a split to its neighbour, unlike fixStartPointIP's
// firstFound flags, :134-143 and :181-196).  Both windows then advance past
// their entries <= m, at least one of them by its whole length.  hit(x, r) is
// called by the lane of every match, r being x's index in the Y window.
// Called by the whole wave with wave-uniform X, nx, Y, ny.
template <class Hit>
__device__ inline uint32_t merge_path_count(const int32_t *__restrict__ X, int32_t nx,
                                            const int32_t *__restrict__ Y, int32_t ny, Hit &&hit)
{
    const int lane = threadIdx.x & (kWave - 1);
    int32_t ix = 0, iy = 0;
    uint32_t h = 0;
    while (ix < nx && iy < ny) {
        const int32_t cx = min(kWave, nx - ix), cy = min(kWave, ny - iy);
        const int32_t x = lane < cx ? X[ix + lane] : INT32_MAX;
        const int32_t y = lane < cy ? Y[iy + lane] : INT32_MAX;
        const int32_t m = min(__builtin_amdgcn_readlane(x, cx - 1), __builtin_amdgcn_readlane(y, cy - 1));
        // r = number of window entries y < x among lanes 0..62 (lane 63 is
        // never probed: if it alone is below x, x is not in the window)
        int32_t r = 0;
#pragma unroll
        for (int st = kWave / 2; st > 0; st >>= 1) {
            const int32_t p = __shfl(y, r + st - 1, kWave);
            r += (p < x) ? st : 0;
        }
        // the shuffle runs on every lane: a lane that ds_bpermute reads from
        // must be active, or the read returns 0
        const int32_t yr = __shfl(y, r, kWave);
        const bool match = x <= m && yr == x;
        h += (uint32_t)__popcll(__ballot(match));
        if (match) hit(x, iy + r);
        ix += __popcll(__ballot(x <= m));
        iy += __popcll(__ballot(y <= m));
    }
    return h;
}

// The part of a task list [pb, pb+pl) inside an owner tile's id range
// [lo_v, hi_v] (tiled owners): two binary searches.
__device__ inline void tile_filter(const int32_t *__restrict__ adj, int64_t &pb, int32_t &pl,
                                   int32_t lo_v, int32_t hi_v)
{
    int32_t lo = 0, hi = pl;
    while (lo < hi) {
        const int32_t mid = (lo + hi) >> 1;
        if (adj[pb + mid] < lo_v) lo = mid + 1; else hi = mid;
    }
    const int32_t a0 = lo;
    hi = pl;
    while (lo < hi) {
        const int32_t mid = (lo + hi) >> 1;
        if (adj[pb + mid] <= hi_v) lo = mid + 1; else hi = mid;
    }
    pb += a0;
    pl = lo - a0;
}

// The small and mid bins run at 7 waves per SIMD (at most 72 VGPRs; the mid
// bin spills a few values to scratch).  Measured at R-MAT 24 and on 4096
// batched graphs, this beats 4 to 6 waves (fewer jobs in flight to hide
// their dependent loads) and, in the mid bin, 8 (64 VGPRs, far more spills:
// 45 -> 65 ms batched).
template <int LMAX, int NW, bool ORIENTED>
#ifndef TC_MID_WAVES
#define TC_MID_WAVES 7
#endif
#ifndef TC_MID_LOOKUPS
#define TC_MID_LOOKUPS 1
#endif
#ifndef TC_SMALL_SLOTS
#define TC_SMALL_SLOTS 2
#endif
#ifndef TC_MID_SLOTS
#define TC_MID_SLOTS 2
#endif
#ifndef TC_LONG_SLOTS
#define TC_LONG_SLOTS 4
#endif
// waves per SIMD the small bin's launch bound asks for (its jobs are short
// chains of dependent loads; 5, 6, 7, 8: 39.0, 34.6, 33.7, 34.0 ms on 4096
// batched graphs)
#ifndef TC_SMALL_WAVES
#define TC_SMALL_WAVES 7
#endif
__global__ __launch_bounds__(NW * kWave, LMAX <= 64 ? TC_SMALL_WAVES : LMAX <= 512 ? TC_MID_WAVES : 1) void k_count(const Job *__restrict__ jobs, int64_t njobs,
                                                      CountArgs A)
{
    // The owner tile lives in a bucketed two-choice hash table (BucketTableT),
    // load <= 1/2, so a lookup reads one bucket, a second only where the
    // first is full, and loops only when both are.  C counts hits per slot
    // (third-vertex counts of the oriented method).
    constexpr int NBMAX = (LMAX + 1) / 2 > 0 ? (LMAX + 1) / 2 : 1;
    __shared__ int4 HB[NBMAX];
    __shared__ uint32_t C[ORIENTED ? 4 * NBMAX : 1];
    __shared__ int32_t pre[NW][kWave + 1];
    __shared__ int64_t tpb[NW][kWave];
    __shared__ int32_t tpl[NW][kWave];
    __shared__ uint32_t tcnt[NW][kWave];
    __shared__ unsigned long long red[NW];
    __shared__ int32_t wsum[NW];
    // bitmap tiles in the mid and largest bins, whose jobs are long enough to
    // amortise building a bitmap over the tile's value range (measured: the
    // mid bin 25.9 -> 23.9 ms at R-MAT 24, 80.7 -> 68.8 ms on 4096 batched
    // graphs; the small bin gains nothing); words and their prefixes share
    // HB's space
    constexpr bool kChunkedLong = LMAX < 4096;
    constexpr bool kUseBitmap = LMAX >= 512;
    constexpr int kBmWords = (16 * NBMAX) / 8;
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = threadIdx.x / kWave;
    const int nthr = NW * kWave;
    unsigned long long hits_all = 0;
    unsigned long long sc[4] = {0, 0, 0, 0};

    // a fixed round-robin deal of jobs: a device counter per job (as in
    // k_count_long's bitmap launch) serialises on one address at these jobs'
    // sizes (small bin 3.3 -> 23 ms at R-MAT 24)
    for (int64_t jb = blockIdx.x; jb < njobs; jb += gridDim.x) {
        const Job J = jobs[jb];
        const VInfo vo = A.vi[J.owner];
        const int64_t obase = vo.rb;
        const int64_t tbase = ORIENTED ? A.toff[J.owner] : vo.rb;
        const int32_t Lfull = ORIENTED ? vo.len - vo.dm : vo.len;
        const int32_t s0 = J.tile * LMAX;
        const int32_t L = min(LMAX, Lfull - s0);
        const bool tiled = LMAX >= 4096 && Lfull > LMAX;  // smaller bins hold whole lists
        const int32_t *Sg = A.adj + obase + s0;  // owner tile (sorted) in global memory
        // 2 ids per bucket in the small and mid bins (an 8-byte read per
        // probe; the small bin 3.8 -> 3.5 ms at R-MAT 24, 36.2 -> 34.8 ms on
        // 4096 batched graphs, the mid bin within noise), 4 in the largest;
        // either way at least 2L slots
        constexpr int kSlots = LMAX <= 64 ? TC_SMALL_SLOTS : LMAX <= 512 ? TC_MID_SLOTS : 4;
        using Table = BucketTableT<kSlots>;
        int nbb = 0;
        while ((kSlots << nbb) < 2 * L) ++nbb;
        const Table T{reinterpret_cast<typename Table::V *>(HB), nbb};
        const int32_t lo_v = Sg[0], hi_v = Sg[L - 1];
        // a tile whose value range is narrow enough gets a bitmap instead of
        // the hash table (same LDS space)
        const uint32_t span = (uint32_t)(hi_v - lo_v) + 1u;
        const int32_t nwords = (int32_t)((span + 31u) >> 5);
        // (one word past the span: the zero guard word)
        const bool use_bm = kUseBitmap && nwords < kBmWords;
        uint2 *BMw = reinterpret_cast<uint2 *>(HB);
        const TileBitmap BT{BMw, lo_v, (uint32_t)nwords << 5};
        if (ORIENTED && !kChunkedLong && A.long_external) {
            // k_count_long takes every hash task: only a job holding a task of
            // another strategy needs its tile here
            bool other = false;
            for (int32_t t = J.t_begin + threadIdx.x; t < J.t_end; t += nthr) {
                const TaskRec tk = A.tasks[tbase + t];
                int32_t pl = tk.len();
                int64_t pb = tk.base();
                if (tiled && pl > 0) tile_filter(A.adj, pb, pl, lo_v, hi_v);
                if (pl > 0 && choose_strategy(L, pl, A) != S_HASH) other = true;
            }
            if (!__syncthreads_or(other)) continue;
        }
        // under the merge-path policy every task walks the sorted tile itself:
        // no table, and third-vertex counters indexed by tile position
        const bool mp_only = A.policy == TC_POLICY_MERGEPATH;
        const bool by_pos = use_bm || mp_only;
        if (mp_only) {
            if (ORIENTED)
                for (int i = threadIdx.x; i < L; i += nthr) C[i] = 0;
            __syncthreads();
        } else if (use_bm) {
            for (int i = threadIdx.x; i <= nwords; i += nthr) BMw[i] = make_uint2(0u, 0u);
            if (ORIENTED)
                for (int i = threadIdx.x; i < L; i += nthr) C[i] = 0;
            __syncthreads();
            for (int i = threadIdx.x; i < L; i += nthr) {
                const uint32_t off = (uint32_t)(Sg[i] - lo_v);
                atomicOr(&BMw[off >> 5].x, 1u << (off & 31u));
            }
            __syncthreads();
            // .y = exclusive prefix of the word popcounts (contiguous chunk per thread)
            const int32_t per = (nwords + nthr - 1) / nthr;
            const int32_t k0 = threadIdx.x * per, k1 = min(nwords, k0 + per);
            int32_t mine = 0;
            for (int32_t k = k0; k < k1; ++k) mine += __popc(BMw[k].x);
            const int32_t incl = wave_incl_scan_i32(mine);
            if (lane == kWave - 1) wsum[wid] = incl;
            __syncthreads();
            int32_t run = incl - mine;
            for (int w = 0; w < wid; ++w) run += wsum[w];
            for (int32_t k = k0; k < k1; ++k) {
                BMw[k].y = (uint32_t)run;
                run += __popc(BMw[k].x);
            }
            __syncthreads();
        } else {
            for (int i = threadIdx.x; i < (1 << nbb); i += nthr) T.HB[i] = Table::W::empty();
            if (ORIENTED)
                for (int i = threadIdx.x; i < kSlots << nbb; i += nthr) C[i] = 0;
            __syncthreads();
            for (int i = threadIdx.x; i < L; i += nthr) T.insert(Sg[i]);
            __syncthreads();
        }
        // batched probes: bitmap reads all issued before use; the hash table one by one
        auto lookup_n = [&](auto &w, auto &slot) {
            constexpr int n = (int)(sizeof(w) / sizeof(w[0]));
            if (use_bm) {
                BT.find_n(w, slot);
            } else {
                // lookups in flight: four in the small bin; one in the mid bin
                // (its spills 112 -> 44 B per lane, bin 19.9 -> 19.4 ms at
                // R-MAT 24); two in the largest bin, where the table is the
                // fallback of the bitmap and registers are scarce
                constexpr int gmax = LMAX >= 4096 ? 2 : (LMAX >= 512 ? TC_MID_LOOKUPS : 4);
                constexpr int g = n < gmax ? n : gmax;
#pragma unroll
                for (int u = 0; u < n; u += g) T.template find_n<g>(&w[u], &slot[u]);
            }
        };
        auto lookup = [&](int32_t w) -> int32_t {
            int32_t ww1[1] = {w}, s1[1];
            lookup_n(ww1, s1);
            return s1[0];
        };
        unsigned long long hits_job = 0;

        // tasks are dealt to the waves round-robin so short jobs still keep every wave busy
        for (int32_t g0 = 0; J.t_begin + (g0 * kWave) * NW + wid < J.t_end; ++g0) {
            const int32_t t = J.t_begin + (g0 * kWave + lane) * NW + wid;
            bool valid = t < J.t_end;
            int32_t tv = 0, sl = 0;
            int64_t sb = 0;
            if (valid) {
                if (ORIENTED) {
                    const TaskRec tk = A.tasks[tbase + t];
                    sl = tk.len();
                    sb = tk.base();
                } else {
                    tv = A.adj[tbase + t];
                }
            }
            VInfo vt = {0, 0, 0};
            if (!ORIENTED && valid) vt = A.vi[tv];
            // oriented: the task list is the suffix of N+(tv) after the owner
            int32_t pl = ORIENTED ? sl : vt.len;
            int64_t pb = ORIENTED ? sb : vt.rb;
            if (!ORIENTED && valid &&
                (tv == J.owner || vt.len < 2 || vo.len < 2 || (A.skip_lower && tv < J.owner)))
                valid = false;
            if (!valid) pl = 0;
            if (tiled && pl > 0) {
                // keep only the part of the task list inside the owner tile's id range
                int32_t lo = 0, hi = pl;
                while (lo < hi) {
                    int32_t mid = (lo + hi) >> 1;
                    if (A.adj[pb + mid] < lo_v) lo = mid + 1; else hi = mid;
                }
                int32_t a0 = lo;
                hi = pl;
                while (lo < hi) {
                    int32_t mid = (lo + hi) >> 1;
                    if (A.adj[pb + mid] <= hi_v) lo = mid + 1; else hi = mid;
                }
                pb += a0;
                pl = lo - a0;
            }
            int strat = (pl > 0) ? choose_strategy(L, pl, A) : -1;
            if (ORIENTED && !kChunkedLong && A.long_external && strat == S_HASH) {
                strat = -1;  // k_count_long's task
                pl = 0;
                valid = false;
            }
            if (strat >= 0) sc[strat]++;
            tpb[wid][lane] = pb;
            tpl[wid][lane] = pl;
            tcnt[wid][lane] = 0;
            uint32_t mcount = 0;
            wave_sync();

            // -- merge path: the whole wave on one task at a time
            for (uint64_t mp = __ballot(strat == S_MPATH); mp; mp &= mp - 1) {
                const int jm = __ffsll((unsigned long long)mp) - 1;
                const int32_t *X = A.adj + readlane_i64(pb, jm);
                const int32_t nx = __builtin_amdgcn_readlane(pl, jm);
                const uint32_t h = merge_path_count(X, nx, Sg, L, [&](int32_t x, int32_t pos) {
                    if (ORIENTED) atomicAdd(&C[by_pos ? pos : lookup(x)], 1u);
                });
                if (lane == jm) tcnt[wid][jm] = h;
            }

            // Long hash tasks are streamed in 64-entry chunks across the whole
            // wave (stream_tasks): one batch of kBigUnroll loads per lane spans
            // several tasks, and hit counts come from ballots.
            // (a bin whose threshold no task reaches compiles no stream)
            constexpr bool kStreams = big_task_min<LMAX>() < (1 << 30);
            const bool is_big = kStreams && strat == S_HASH && pl >= big_task_min<LMAX>();
            if (!kChunkedLong) {
                // largest bin: one long task at a time across the wave
                uint64_t big = __ballot(is_big);
                while (big) {
                    const int jb2 = __ffsll((unsigned long long)big) - 1;
                    big &= big - 1;
                    const int32_t *lst = A.adj + readlane_i64(pb, jb2);
                    const int32_t n = __builtin_amdgcn_readlane(pl, jb2);
                    uint32_t h = 0;
                    for (int32_t k0 = lane; k0 < n; k0 += kBigUnroll * kWave) {
                        int32_t w[kBigUnroll];
#pragma unroll
                        for (int u = 0; u < kBigUnroll; ++u) {
                            // masked loads: a slot past the end of the task is
                            // skipped by the whole wave
                            const int32_t k = k0 + u * kWave;
                            w[u] = (k < n) ? lst[k] : -1;
                        }
                        int32_t slots[kBigUnroll];
                        lookup_n(w, slots);
#pragma unroll
                        for (int u = 0; u < kBigUnroll; ++u) {
                            if (slots[u] >= 0) {
                                ++h;
                                if (ORIENTED) atomicAdd(&C[slots[u]], 1u);
                            }
                        }
                    }
                    h = wave_sum<uint32_t>(h);
                    if (lane == jb2) tcnt[wid][jb2] = h;
                }
            } else if (kStreams && __ballot(is_big)) {
                const uint32_t h = stream_tasks<kBigUnroll, ORIENTED>(
                    A.adj, is_big ? pl : 0, pb, [&](const int32_t (&w)[kBigUnroll], uint32_t (&hc)[kBigUnroll]) {
                        int32_t slots[kBigUnroll];
                        lookup_n(w, slots);
#pragma unroll
                        for (int u = 0; u < kBigUnroll; ++u) {
                            const bool hit = slots[u] >= 0;
                            hc[u] = (uint32_t)__popcll(__ballot(hit));
                            if (ORIENTED && hit) atomicAdd(&C[slots[u]], 1u);
                        }
                    });
                if (is_big) mcount = h;
            }

            // Flattened phases over the remaining tasks of this wave: items of
            // the hash phase are task-list entries, items of the search phase
            // are owner-tile entries.  Each lane keeps a running task index and
            // a private hit count, flushed when its task changes.
#pragma unroll 1
            for (int phase = 0; phase < 2; ++phase) {
                const int want = phase == 0 ? S_HASH : S_BSEARCH;
                const int32_t items =
                    (strat == want && !(phase == 0 && kStreams && pl >= big_task_min<LMAX>())) ? (phase == 0 ? pl : L) : 0;
                const int32_t incl = wave_incl_scan_i32(items);
                wave_sync();
                pre[wid][lane + 1] = incl;
                if (lane == 0) pre[wid][0] = 0;
                const int32_t total = __shfl(incl, kWave - 1, kWave);
                wave_sync();
                if (total == 0) continue;
                int32_t j = 0, myj = -1;
                uint32_t myhits = 0;
                for (int32_t base = lane; base < total; base += kUnroll * kWave) {
                    int32_t jj[kUnroll], ww[kUnroll], kk[kUnroll];
                    bool ok[kUnroll];
#pragma unroll
                    for (int u = 0; u < kUnroll; ++u) {
                        const int32_t it = base + u * kWave;
                        ok[u] = it < total;
                        jj[u] = j;
                        kk[u] = 0;
                        const int32_t *src = A.adj;  // a valid address for lanes past the end
                        if (ok[u]) {
                            if (pre[wid][j + 1] <= it) {
                                int32_t lo = j + 1, hi = kWave;  // largest j with pre[j] <= it
                                while (hi - lo > 1) {
                                    int32_t mid = (lo + hi) >> 1;
                                    if (pre[wid][mid] <= it) lo = mid; else hi = mid;
                                }
                                j = lo;
                            }
                            jj[u] = j;
                            kk[u] = it - pre[wid][j];
                            src = (phase == 0) ? A.adj + tpb[wid][j] + kk[u] : Sg + kk[u];
                        }
                        // the load sits outside the branch: loads under an exec
                        // mask make the compiler wait for all of them at once
                        const int32_t x = *src;
                        ww[u] = ok[u] ? x : -1;
                    }
                    if (phase == 0) {
                        int32_t slots[kUnroll];
                        lookup_n(ww, slots);
#pragma unroll
                        for (int u = 0; u < kUnroll; ++u) {
                            const int32_t slot = slots[u];
                            if (!ok[u]) continue;
                            if (jj[u] != myj) {
                                if (myhits) atomicAdd(&tcnt[wid][myj], myhits);
                                myj = jj[u];
                                myhits = 0;
                            }
                            if (slot >= 0) {
                                ++myhits;
                                if (ORIENTED) atomicAdd(&C[slot], 1u);
                            }
                        }
                    } else {
                        // interleaved lower_bound searches of owner entries in task lists
                        const int32_t *lst[kUnroll];
                        int32_t b0[kUnroll], n0[kUnroll], nn[kUnroll];
                        int32_t steps = 0;
#pragma unroll
                        for (int u = 0; u < kUnroll; ++u) {
                            lst[u] = A.adj + tpb[wid][jj[u]];
                            n0[u] = ok[u] ? tpl[wid][jj[u]] : 0;
                            nn[u] = n0[u];
                            b0[u] = 0;
                            steps = max(steps, 32 - __clz(n0[u]));
                        }
                        for (int32_t s = 0; s < steps; ++s) {
#pragma unroll
                            for (int u = 0; u < kUnroll; ++u) {
                                if (nn[u] > 0) {
                                    const int32_t half = nn[u] >> 1;
                                    if (lst[u][b0[u] + half] < ww[u]) {
                                        b0[u] += half + 1;
                                        nn[u] -= half + 1;
                                    } else {
                                        nn[u] = half;
                                    }
                                }
                            }
                        }
#pragma unroll
                        for (int u = 0; u < kUnroll; ++u) {
                            if (!ok[u]) continue;
                            if (jj[u] != myj) {
                                if (myhits) atomicAdd(&tcnt[wid][myj], myhits);
                                myj = jj[u];
                                myhits = 0;
                            }
                            if (b0[u] < n0[u] && lst[u][b0[u]] == ww[u]) {
                                ++myhits;
                                if (ORIENTED) atomicAdd(&C[lookup(ww[u])], 1u);
                            }
                        }
                    }
                }
                if (myhits) atomicAdd(&tcnt[wid][myj], myhits);
            }
            // -- merge: one lane walks both sorted lists
            if (strat == S_MERGE) {
                const int32_t *lst = A.adj + pb;
                int32_t i = 0, k = 0;
                {
                    int32_t first = lst[0], a = 0, b = L;
                    while (a < b) {
                        int32_t mid = (a + b) >> 1;
                        if (Sg[mid] < first) a = mid + 1; else b = mid;
                    }
                    k = a;
                }
                while (i < pl && k < L) {
                    int32_t x = lst[i], y = Sg[k];
                    if (x == y) {
                        ++mcount;
                        if (ORIENTED) atomicAdd(&C[lookup(x)], 1u);
                    }
                    i += (x <= y);
                    k += (x >= y);
                }
            }
            wave_sync();
            const uint32_t c = tcnt[wid][lane] + mcount;
            if (valid && c) {
                hits_job += c;
                // the task's hits go to its N+ entry (the one before its suffix),
                // summed per vertex after the count: a store, not a global atomic
                // (a tiled owner meets the task once per tile)
                if (ORIENTED) {
                    if (tiled)
                        atomicAdd(&A.pcnt[sb - 1], c);
                    else
                        A.pcnt[sb - 1] = c;
                }
            }
            wave_sync();
        }
        // owner + third-vertex flush
        unsigned long long hj = wave_sum<unsigned long long>(hits_job);
        if (lane == 0) red[wid] = hj;
        __syncthreads();
        unsigned long long hjob = 0;
        for (int w = 0; w < NW; ++w) hjob += red[w];
        if (ORIENTED) {
            if (by_pos) {
                for (int i = threadIdx.x; i < L; i += nthr) {
                    const uint32_t ci = C[i];
                    if (ci) atomicAdd(&A.out[A.order[Sg[i]]], 2u * ci);
                }
            } else {
                for (int i = threadIdx.x; i < kSlots << nbb; i += nthr) {
                    const uint32_t ci = C[i];
                    if (ci) atomicAdd(&A.out[A.order[T.id_at(i)]], 2u * ci);
                }
            }
        }
        const int32_t owner_v = ORIENTED ? A.order[J.owner] : J.owner;
        if (threadIdx.x == 0 && hjob)
            atomicAdd(&A.out[owner_v], ORIENTED ? (uint32_t)(2 * hjob) : (uint32_t)hjob);
        hits_all += (threadIdx.x == 0) ? hjob : 0ull;
        if (A.per_graph && threadIdx.x == 0 && hjob)
            atomicAdd(&A.per_graph[A.gid[J.owner]], (ORIENTED ? 6ull : 1ull) * hjob);
        __syncthreads();
    }
    if (threadIdx.x == 0 && hits_all) atomicAdd(A.global, (ORIENTED ? 6ull : 1ull) * hits_all);
    if (A.strat) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            unsigned long long s = wave_sum<unsigned long long>(sc[q]);
            if (lane == 0 && s) atomicAdd(q < 3 ? &A.strat[q] : A.strat_mp, s);
        }
    }
}

// Hash-strategy tasks of the largest bin, oriented method (almost all of the
// bin's items).  A wave's 64 tasks are streamed by stream_tasks as one
// sequence of 64-entry chunks, kLongUnroll chunks per batch, with the next
// batch's loads in flight while the current one is probed.  BM = true walks
// every job, takes those whose owner tile fits the bitmap (it needs only the
// bitmap and L counters in LDS: three blocks per CU) and lists the others for
// BM = false, which takes the hash-table and wide-bitmap tiles from that list;
// k_count<LMAX, NW, true> with long_external does the tasks of the other
// strategies.
#ifndef TC_LONG_UNROLL
#define TC_LONG_UNROLL 8
#endif
constexpr int kLongUnroll = TC_LONG_UNROLL;

template <int LMAX, int NW, bool BM>
__global__ __launch_bounds__(NW * kWave, BM ? 6 : 4) void k_count_long(const Job *__restrict__ jobs,
                                                           int64_t njobs, CountArgs A)
{
    constexpr int NBMAX = LMAX / 2;
    constexpr int kBmWords = 2 * NBMAX;  // uint2 words in HB's space
    // one LDS block: the table (or bitmap) HB, then the counters C.  The
    // hash-table launch reuses all of it for a wider bitmap when a tile's
    // span allows: kWideWords words, then L counters
    constexpr int kRawInt4 = BM ? NBMAX + LMAX / 4 : 2 * NBMAX;
    constexpr int kWideWords = BM ? 0 : (kRawInt4 * 16 - LMAX * 4) / 8;
    __shared__ int4 RAW[kRawInt4];
    int4 *HB = RAW;
    uint32_t *C = reinterpret_cast<uint32_t *>(RAW + NBMAX);
    __shared__ unsigned long long red[NW];
    __shared__ int32_t wsum[NW];
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = threadIdx.x / kWave;
    const int nthr = NW * kWave;
    unsigned long long hits_all = 0, n_hash = 0, n_items = 0, n_tile = 0;

    // Jobs come from device counters, so blocks that drew cheap jobs take
    // more and a launch ends together (a fixed round-robin deal left a long
    // tail: 105 -> 88 ms at R-MAT 24).  The bitmap launch walks every job and
    // lists the ones it leaves to the hash-table launch, which walks that list.
    __shared__ int64_t s_job;
    int64_t jnext = blockIdx.x;
    for (;;) {
        int64_t jb = jnext;
        if (BM || A.hjobs) {
            if (threadIdx.x == 0) {
                if (BM) {
                    s_job = (int64_t)atomicAdd(A.job_ctr, 1ull);
                } else {
                    const int64_t k = (int64_t)atomicAdd(A.job_ctr + 2, 1ull);
                    s_job = k < (int64_t)A.job_ctr[1] ? A.hjobs[k] : njobs;
                }
            }
            __syncthreads();
            jb = s_job;
            __syncthreads();
        } else {
            jnext += gridDim.x;
        }
        if (jb >= njobs) break;
        const Job J = jobs[jb];
        const VInfo vo = A.vi[J.owner];
        const int64_t tbase = A.toff[J.owner];
        const int32_t Lfull = vo.len - vo.dm;
        const int32_t s0 = J.tile * LMAX;
        const int32_t L = min(LMAX, Lfull - s0);
        const bool tiled = LMAX >= 4096 && Lfull > LMAX;  // smaller bins hold whole lists
        const int32_t *Sg = A.adj + vo.rb + s0;
        const int32_t lo_v = Sg[0], hi_v = Sg[L - 1];
        const uint32_t span = (uint32_t)(hi_v - lo_v) + 1u;
        const int32_t nwords = (int32_t)((span + 31u) >> 5);
        // (a tile needs one word past its span: the zero guard word)
        if ((nwords < kBmWords) != BM) {  // the other instantiation's job
            if (BM && A.hjobs && threadIdx.x == 0) A.hjobs[atomicAdd(A.job_ctr + 1, 1ull)] = jb;
            continue;
        }
        using Table = BucketTableT<TC_LONG_SLOTS>;
        constexpr int kSlots = TC_LONG_SLOTS;
        int nbb = 0;
        while ((kSlots << nbb) < 2 * L) ++nbb;
        const Table T{reinterpret_cast<typename Table::V *>(HB), nbb};
        uint2 *BMw = reinterpret_cast<uint2 *>(RAW);
        // bitmap: always in the bitmap launch, and in the hash-table launch
        // when the span fits the wide layout (counters after the words)
        const bool use_bm = BM || nwords < kWideWords;
        uint32_t *Cb = BM ? C : reinterpret_cast<uint32_t *>(RAW) + 2 * kWideWords;
        if (use_bm) {
            for (int i = threadIdx.x; i <= nwords; i += nthr) BMw[i] = make_uint2(0u, 0u);
            for (int i = threadIdx.x; i < L; i += nthr) Cb[i] = 0;
            __syncthreads();
            for (int i = threadIdx.x; i < L; i += nthr) {
                const uint32_t off = (uint32_t)(Sg[i] - lo_v);
                atomicOr(&BMw[off >> 5].x, 1u << (off & 31u));
            }
            __syncthreads();
            const int32_t per = (nwords + nthr - 1) / nthr;
            const int32_t k0 = threadIdx.x * per, k1 = min(nwords, k0 + per);
            int32_t mine = 0;
            for (int32_t k = k0; k < k1; ++k) mine += __popc(BMw[k].x);
            const int32_t incl = wave_incl_scan_i32(mine);
            if (lane == kWave - 1) wsum[wid] = incl;
            __syncthreads();
            int32_t run = incl - mine;
            for (int w = 0; w < wid; ++w) run += wsum[w];
            for (int32_t k = k0; k < k1; ++k) {
                BMw[k].y = (uint32_t)run * 4u;
                run += __popc(BMw[k].x);
            }
            __syncthreads();
        } else {
            for (int i = threadIdx.x; i < (1 << nbb); i += nthr) T.HB[i] = Table::W::empty();
            for (int i = threadIdx.x; i < kSlots << nbb; i += nthr) C[i] = 0;
            __syncthreads();
            for (int i = threadIdx.x; i < L; i += nthr) T.insert(Sg[i]);
            __syncthreads();
        }
        uint32_t *Ct = use_bm ? Cb : C;  // hit counters by slot
        // offsets past the span are clamped to the zero word at index nwords
        const uint32_t guard = (uint32_t)nwords << 5;
        unsigned long long hits_job = 0;

        // the same round-robin deal of tasks to waves as k_count
        for (int32_t g0 = 0; J.t_begin + (g0 * kWave) * NW + wid < J.t_end; ++g0) {
            const int32_t t = J.t_begin + (g0 * kWave + lane) * NW + wid;
            int32_t pl = 0;
            int64_t pb = 0, sb = 0;
            if (t < J.t_end) {
                const TaskRec tk = A.tasks[tbase + t];
                pl = tk.len();
                sb = tk.base();
                pb = sb;
            }
            if (tiled && pl > 0) tile_filter(A.adj, pb, pl, lo_v, hi_v);
            if (!(pl > 0 && choose_strategy(L, pl, A) == S_HASH)) pl = 0;
            n_hash += pl > 0;
            n_items += (unsigned long long)pl;
            // 0 never hits (owner lists hold ranks above their owner): it
            // clamps to the zero guard word, and the hash table holds no 0
            const uint32_t mine_h = stream_tasks<kLongUnroll, true>(A.adj, pl, pb, [&](const int32_t (&w)[kLongUnroll], uint32_t (&hc)[kLongUnroll]) {
                if (use_bm) {
                    uint32_t off[kLongUnroll];
                    uint2 x[kLongUnroll];
#pragma unroll
                    for (int u = 0; u < kLongUnroll; ++u) {
                        off[u] = min((uint32_t)(w[u] - lo_v), guard);  // w < lo wraps high
                        x[u] = BMw[off[u] >> 5];
                    }
#pragma unroll
                    for (int u = 0; u < kLongUnroll; ++u) {
                        // v_bfe_u32 reads bits 4:0 of its offset and width; the
                        // word's prefix is kept in bytes (counter index * 4)
                        const bool hit = __builtin_amdgcn_ubfe(x[u].x, off[u], 1u) != 0u;
                        hc[u] = (uint32_t)__popcll(__ballot(hit));
                        if (hit) {
                            const uint32_t byte = x[u].y + 4u * __popc(__builtin_amdgcn_ubfe(x[u].x, 0u, off[u]));
                            atomicAdd(reinterpret_cast<uint32_t *>(reinterpret_cast<char *>(Ct) + byte), 1u);
                        }
                    }
                } else {
                    int32_t slots[kLongUnroll];
#pragma unroll
                    for (int u = 0; u < kLongUnroll; u += 2) T.template find_n<2>(&w[u], &slots[u]);
#pragma unroll
                    for (int u = 0; u < kLongUnroll; ++u) {
                        const bool hit = slots[u] >= 0;
                        hc[u] = (uint32_t)__popcll(__ballot(hit));
                        if (hit) atomicAdd(&Ct[slots[u]], 1u);
                    }
                }
            });
            if (mine_h) {
                hits_job += mine_h;
                // a tiled owner meets the task once per tile
                if (tiled)
                    atomicAdd(&A.pcnt[sb - 1], mine_h);
                else
                    A.pcnt[sb - 1] = mine_h;
            }
        }
        // owner + third-vertex flush
        const unsigned long long hj = wave_sum<unsigned long long>(hits_job);
        if (lane == 0) red[wid] = hj;
        __syncthreads();
        unsigned long long hjob = 0;
        for (int w = 0; w < NW; ++w) hjob += red[w];
        if (use_bm) {
            for (int i = threadIdx.x; i < L; i += nthr) {
                const uint32_t ci = Cb[i];
                if (ci) atomicAdd(&A.outr[Sg[i]], 2u * ci);
            }
        } else {
            for (int i = threadIdx.x; i < kSlots << nbb; i += nthr) {
                const uint32_t ci = C[i];
                if (ci) atomicAdd(&A.outr[T.id_at(i)], 2u * ci);
            }
        }
        if (threadIdx.x == 0 && hjob) {
            atomicAdd(&A.outr[J.owner], (uint32_t)(2 * hjob));
            if (A.per_graph) atomicAdd(&A.per_graph[A.gid[J.owner]], 6ull * hjob);
        }
        hits_all += (threadIdx.x == 0) ? hjob : 0ull;
        n_tile += (threadIdx.x == 0) ? (unsigned long long)L : 0ull;
        __syncthreads();
    }
    if (threadIdx.x == 0 && hits_all) atomicAdd(A.global, 6ull * hits_all);
    if (A.strat) {
        const unsigned long long nh = wave_sum<unsigned long long>(n_hash);
        if (lane == 0 && nh) atomicAdd(&A.strat[S_HASH], nh);
    }
    if (A.long_stats) {
        const unsigned long long v[3] = {n_hash, n_items, n_tile};
        block_add_u64<3>(v, A.long_stats + (BM ? 0 : 3));
    }
}

// ------------------------------------------------------------- driver --

tc_schedule default_schedule()
{
    tc_schedule s = {};
    s.method = TC_METHOD_AUTO;
    s.policy = TC_POLICY_FIXED;
    s.small_max = 64;
    s.mid_max = 512;
    s.job_items = 8192;
    s.ratio_x16 = 0;   // no merges by the fixed heuristic (64: long <= 4x short, <= 96)
    s.persistent = 1;
    s.flush_items = 1024;
    return s;
}

namespace {

constexpr int kL0 = 64, kL1 = 512, kL2 = 4096;
// owners of at least this degree get their task lists from their own rows
// (k_orient_transpose_hubs) instead of by appends
constexpr int kHubDegree = 65536;  // with the appends
#ifndef TC_MID_NW
#define TC_MID_NW 2
#endif
constexpr int kNW0 = 1, kNW1 = TC_MID_NW, kNW2 = 8;

// ev_long (bin 2, oriented): events recorded before and after the bitmap-tile
// launch of k_count_long and after its hash-table-tile launch.
int launch_bin(int bin, bool oriented, const Job *jobs, int64_t njobs, const CountArgs &A,
               hipStream_t s, hipEvent_t *ev_long)
{
    if (njobs <= 0) return TC_OK;
    int dev = 0;
    TC_HIP(hipGetDevice(&dev));
    int cus = 0;
    TC_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    // persistent grid: enough resident blocks to fill every CU
    int per_cu = bin == 0 ? 4 * TC_SMALL_WAVES / kNW0 : (bin == 1 ? 4 * TC_MID_WAVES / kNW1 : 2);
    int64_t grid = std::min<int64_t>(njobs, (int64_t)cus * per_cu);
    if (bin == 0) {
        if (oriented)
            hipLaunchKernelGGL((k_count<kL0, kNW0, true>), dim3((unsigned)grid), dim3(kNW0 * kWave), 0, s, jobs, njobs, A);
        else
            hipLaunchKernelGGL((k_count<kL0, kNW0, false>), dim3((unsigned)grid), dim3(kNW0 * kWave), 0, s, jobs, njobs, A);
    } else if (bin == 1) {
        if (oriented)
            hipLaunchKernelGGL((k_count<kL1, kNW1, true>), dim3((unsigned)grid), dim3(kNW1 * kWave), 0, s, jobs, njobs, A);
        else
            hipLaunchKernelGGL((k_count<kL1, kNW1, false>), dim3((unsigned)grid), dim3(kNW1 * kWave), 0, s, jobs, njobs, A);
    } else {
        if (oriented) {
            // short tasks here, long hash tasks in k_count_long (bitmap and
            // hash-table tiles in separate launches for their LDS footprints)
            // (the strategy rules must be the same in all three launches; with
            // every task hashed, k_count has nothing left to do)
            CountArgs Al = A;
            Al.hash_only = 1;
            CountArgs As = Al;
            As.long_external = 1;
            // (forced non-hash policies leave k_count_long nothing to stream)
            const bool forced = A.policy == TC_POLICY_MERGE || A.policy == TC_POLICY_BSEARCH ||
                                A.policy == TC_POLICY_MERGEPATH;
            if (forced)
                hipLaunchKernelGGL((k_count<kL2, kNW2, true>), dim3((unsigned)grid), dim3(kNW2 * kWave), 0, s, jobs, njobs, As);
            const int64_t gb = std::min<int64_t>(njobs, (int64_t)cus * 3);
            const int64_t gh = std::min<int64_t>(njobs, (int64_t)cus * 2);
            if (ev_long) TC_HIP(hipEventRecord(ev_long[0], s));
            if (!forced)
                hipLaunchKernelGGL((k_count_long<kL2, kNW2, true>), dim3((unsigned)gb), dim3(kNW2 * kWave), 0, s, jobs, njobs, Al);
            if (ev_long) TC_HIP(hipEventRecord(ev_long[1], s));
            if (!forced)
                hipLaunchKernelGGL((k_count_long<kL2, kNW2, false>), dim3((unsigned)gh), dim3(kNW2 * kWave), 0, s, jobs, njobs, Al);
            if (ev_long) TC_HIP(hipEventRecord(ev_long[2], s));
        } else
            hipLaunchKernelGGL((k_count<kL2, kNW2, false>), dim3((unsigned)grid), dim3(kNW2 * kWave), 0, s, jobs, njobs, A);
    }
    TC_HIP(hipGetLastError());
    return TC_OK;
}

}  // namespace

int build_direct_vinfo(Graph &g)
{
    const int32_t nv = g.st.nv;
    TC_TRY(g.vinfo.ensure((size_t)std::max<int32_t>(nv, 1) * sizeof(VInfo)));
    hipLaunchKernelGGL(k_direct_vinfo, dim3(grid_for(nv, 256, 65536)), dim3(256), 0, g.stream,
                       g.st.boff.as<int64_t>(), g.st.len.as<int32_t>(), nv, g.vinfo.as<VInfo>());
    TC_HIP(hipGetLastError());
    return TC_OK;
}

int prepare_views(Graph &g, bool oriented, int32_t hub_degree, int fallback)
{
    const int32_t nv = g.st.nv;
    const int64_t nnz = g.st.nnz;
    hipStream_t s = g.stream;
    const int64_t nchunks = ceil_div(nnz, kWave);
    const int wblocks = grid_for(nchunks, 4, 1 << 16);  // 4 waves of 256 threads
    const int vblocks = grid_for(nv, 256, 65536);
    g.oriented_views = false;
    if (!oriented) {
        // compact offsets of the store rows, chunk rows, views of the rows as stored
        TC_TRY(g.coff.ensure((size_t)(nv + 1) * sizeof(int64_t)));
        int64_t *coff = g.coff.as<int64_t>();
        TC_TRY(scan_counts_i32(g.st.len.as<int32_t>(), coff, nv, g.tmp, s));
        TC_TRY(g.crow.ensure((size_t)(nchunks + 1) * sizeof(int32_t)));
        hipLaunchKernelGGL(k_chunk_rows, dim3(vblocks), dim3(256), 0, s, coff, nv,
                           g.crow.as<int32_t>());
        TC_HIP(hipGetLastError());
        return build_direct_vinfo(g);
    }

    // ---- store-order offsets and chunk rows
    TC_TRY(g.ov[0].ensure((size_t)(nv + 1) * sizeof(int64_t)));      // store-order offsets
    TC_TRY(g.ov[1].ensure((size_t)(nchunks + 1) * sizeof(int32_t))); // their chunk rows
    int64_t *ocoff = g.ov[0].as<int64_t>();
    int32_t *ocrow = g.ov[1].as<int32_t>();
    TC_TRY(scan_counts_i32(g.st.len.as<int32_t>(), ocoff, nv, g.tmp, s));
    hipLaunchKernelGGL(k_chunk_rows, dim3(vblocks), dim3(256), 0, s, ocoff, nv, ocrow);
    // ---- rank order: stable sort of the degrees
    TC_TRY(g.order.ensure((size_t)nv * sizeof(int32_t)));
    TC_TRY(g.rnk.ensure((size_t)nv * sizeof(int32_t)));
    TC_TRY(g.ov[3].ensure((size_t)nv * sizeof(uint64_t)));
    int32_t *order = g.order.as<int32_t>();
    int32_t *rnk = g.rnk.as<int32_t>();
    uint64_t *keys = g.ov[3].as<uint64_t>();
    // device scalars: [0] maximum degree, [1] first hub rank, [2] plus entries
    // that point to a hub
    TC_TRY(g.ov[12].ensure(4 * sizeof(unsigned long long)));
    unsigned long long *sc = g.ov[12].as<unsigned long long>();
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, s, sc, (int64_t)4);
    hipLaunchKernelGGL(k_rank_keys, dim3(grid_for(nv, 256, 1024)), dim3(256), 0, s,
                       g.st.len.as<int32_t>(), nv, keys, order, sc);
    TC_HIP(hipGetLastError());
    // sort only the bits the current maximum degree needs (one pass per 8)
    unsigned long long hmax = g.plan.hmax;
    if (!g.capturing) {
        TC_HIP(hipMemcpyAsync(&hmax, sc, sizeof(hmax), hipMemcpyDeviceToHost, s));
        TC_HIP(hipStreamSynchronize(s));
        g.plan.hmax = hmax;
    }
    int kbits = std::max(1, ceil_log2_i((int64_t)hmax + 1));
    if (g.batch_voff && g.batch_ngraphs > 1) {
        hipLaunchKernelGGL(k_rank_keys_graph, dim3(vblocks), dim3(256), 0, s, g.batch_voff,
                           g.batch_ngraphs, nv, kbits, keys);
        TC_HIP(hipGetLastError());
        kbits += ceil_log2_i((int64_t)g.batch_ngraphs);
    }
    TC_TRY(radix_sort_u64(keys, order, nv, kbits, g.ov[4], g.ov[5], g.tmp, s));
    hipLaunchKernelGGL(k_rank_inverse, dim3(vblocks), dim3(256), 0, s, order, nv, rnk);
    TC_HIP(hipGetLastError());


    // ---- store-order pass: which entries point to a higher-rank vertex,
    // by rank, and every entry's neighbour as a rank for the N+ write.  If
    // the 4 bytes per entry for those ranks cannot be allocated, the mask
    // compares degrees and the N+ write looks the ranks up again.
    // scratch: mask (u64 x nchunks), cscan (i64 x nchunks+1), cnt (i32 x nchunks)
    TC_TRY(g.tmp2.ensure((size_t)nchunks * 8 + (size_t)(nchunks + 1) * 8 + (size_t)nchunks * 4 + 64));
    uint64_t *mask = g.tmp2.as<uint64_t>();
    int64_t *cscan = reinterpret_cast<int64_t *>(mask + nchunks);
    int32_t *cnt = reinterpret_cast<int32_t *>(cscan + nchunks + 1);
    const bool by_rank =
        !fallback && g.ov[18].ensure((size_t)std::max<int64_t>(nnz, 1) * sizeof(int32_t)) == TC_OK;
    int32_t *tr = by_rank ? g.ov[18].as<int32_t>() : nullptr;
    if (by_rank)
        hipLaunchKernelGGL(k_orient_mask_rank, dim3(wblocks), dim3(256), 0, s, ocoff, ocrow,
                           g.st.boff.as<int64_t>(), g.st.pool.as<int32_t>(), rnk, nv, nnz, mask, cnt,
                           tr);
    else
        hipLaunchKernelGGL(k_orient_mask, dim3(wblocks), dim3(256), 0, s, ocoff, ocrow,
                           g.st.boff.as<int64_t>(), g.st.len.as<int32_t>(), g.st.pool.as<int32_t>(),
                           nv, nnz, mask, cnt);
    TC_HIP(hipGetLastError());
    TC_TRY(scan_counts_i32(cnt, cscan, nchunks, g.tmp, s));
    TC_TRY(g.ov[2].ensure((size_t)nv * sizeof(int32_t)));  // |N-(v)| in store order
    int32_t *dmo = g.ov[2].as<int32_t>();
    hipLaunchKernelGGL(k_orient_dm, dim3(vblocks), dim3(256), 0, s, ocoff, g.st.len.as<int32_t>(),
                       cscan, mask, nv, dmo);
    TC_HIP(hipGetLastError());
    int32_t *dpr = reinterpret_cast<int32_t *>(keys);  // keys are no longer needed
    int32_t *dmr = dpr + nv;
    hipLaunchKernelGGL(k_rank_of, dim3(vblocks), dim3(256), 0, s, order, g.st.len.as<int32_t>(), dmo,
                       nv, rnk, dpr, dmr);
    TC_HIP(hipGetLastError());

    // ---- rank layout: plus array offsets (pcoff), task offsets (coff), views
    TC_TRY(g.ov[6].ensure((size_t)(nv + 1) * sizeof(int64_t)));
    int64_t *pcoff = g.ov[6].as<int64_t>();
    TC_TRY(scan_counts_i32(dpr, pcoff, nv, g.tmp, s));
    TC_TRY(g.coff.ensure((size_t)(nv + 1) * sizeof(int64_t)));
    int64_t *toff = g.coff.as<int64_t>();
    TC_TRY(scan_counts_i32(dmr, toff, nv, g.tmp, s));
    TC_TRY(g.vinfo.ensure((size_t)nv * sizeof(VInfo)));
    VInfo *vi = g.vinfo.as<VInfo>();
    hipLaunchKernelGGL(k_rank_vinfo, dim3(vblocks), dim3(256), 0, s, pcoff, dpr, dmr, nv, vi);
    TC_HIP(hipGetLastError());
    int64_t tot[2] = {g.plan.nplus, g.plan.ntasks};
    if (!g.capturing) {
        TC_HIP(hipMemcpyAsync(&tot[0], pcoff + nv, sizeof(int64_t), hipMemcpyDeviceToHost, s));
        TC_HIP(hipMemcpyAsync(&tot[1], toff + nv, sizeof(int64_t), hipMemcpyDeviceToHost, s));
        TC_HIP(hipStreamSynchronize(s));
        g.plan.nplus = tot[0];
        g.plan.ntasks = tot[1];
    }
    const int64_t nplus = tot[0], ntasks = tot[1];
    TC_TRY(g.padj.ensure((size_t)std::max<int64_t>(nplus, 1) * sizeof(int32_t)));
    TC_TRY(g.tasks.ensure((size_t)std::max<int64_t>(ntasks, 1) * sizeof(TaskRec)));
    int32_t *padj = g.padj.as<int32_t>();
    if (by_rank) {
        int64_t *pdst = reinterpret_cast<int64_t *>(keys);  // (nv words; dpr/dmr are read by now)
        hipLaunchKernelGGL(k_plus_dst, dim3(vblocks), dim3(256), 0, s, ocoff, cscan, mask, rnk, pcoff,
                           nv, pdst);
        hipLaunchKernelGGL(k_orient_write_plus_rank, dim3(wblocks), dim3(256), 0, s, ocoff, ocrow, tr,
                           cscan, mask, pdst, nv, nnz, padj);
    } else
        hipLaunchKernelGGL(k_orient_write_plus, dim3(wblocks), dim3(256), 0, s, ocoff, ocrow,
                           g.st.boff.as<int64_t>(), g.st.pool.as<int32_t>(), cscan, mask, rnk, pcoff,
                           nv, nnz, padj);
    TC_HIP(hipGetLastError());

    // ---- sort every N+ part by rank
    const int64_t cap[5] = {std::max<int64_t>(1, std::min<int64_t>(nv, nplus / 9 + 1)),
                            std::max<int64_t>(1, nplus / 65 + 1),
                            std::max<int64_t>(1, nplus / 513 + 1),
                            std::max<int64_t>(1, nplus / 4097 + 1),
                            std::max<int64_t>(1, nplus / 16385 + 1)};
    int64_t hoff[5];
    int64_t lst_total = 0;
    for (int q = 0; q < 5; ++q) {
        hoff[q] = lst_total;
        lst_total += cap[q];
    }
    TC_TRY(g.ov[7].ensure((size_t)lst_total * sizeof(int32_t)));
    TC_TRY(g.ov[8].ensure(5 * sizeof(int64_t) + 8 * sizeof(int32_t)));
    int32_t *lst = g.ov[7].as<int32_t>();
    int64_t *loff = g.ov[8].as<int64_t>();
    int32_t *lcnt = reinterpret_cast<int32_t *>(loff + 5);  // [5] counts, [5] symmetry flag
    // list offsets set by a kernel (a copy from host memory could not be
    // recorded into a hipGraph)
    hipLaunchKernelGGL(k_set_list_offsets, dim3(1), dim3(64), 0, s, loff,
                       ListOffsets{{hoff[0], hoff[1], hoff[2], hoff[3], hoff[4]}},
                       reinterpret_cast<uint32_t *>(lcnt));
    hipLaunchKernelGGL(k_sort_plus_small, dim3(vblocks), dim3(256), 0, s, vi, nv, padj, lst, loff,
                       lcnt);
    TC_HIP(hipGetLastError());
    int dev = 0, cus = 0;
    TC_HIP(hipGetDevice(&dev));
    TC_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    hipLaunchKernelGGL(k_sort_plus_wave, dim3(cus * 8), dim3(256), 0, s, vi, lst + hoff[0], lcnt + 0,
                       padj);
    hipLaunchKernelGGL((k_sort_plus_block<512, 64>), dim3(cus * TC_SORT512_BLOCKS), dim3(64), 0, s,
                       vi, lst + hoff[1], lcnt + 1, padj);
    hipLaunchKernelGGL((k_sort_plus_block<4096, 256>), dim3(cus * TC_SORT4096_BLOCKS), dim3(256), 0, s,
                       vi, lst + hoff[2], lcnt + 2, padj);
    hipLaunchKernelGGL((k_sort_plus_block<16384, 1024>), dim3(cus), dim3(1024), 0, s, vi,
                       lst + hoff[3], lcnt + 3, padj);
    TC_HIP(hipGetLastError());
    if (!g.capturing) {
        int32_t nhuge = 0;
        TC_HIP(hipMemcpyAsync(&nhuge, lcnt + 4, sizeof(int32_t), hipMemcpyDeviceToHost, s));
        TC_HIP(hipStreamSynchronize(s));
        g.plan.huge.clear();
        if (nhuge > 0) {
            std::vector<int32_t> rows((size_t)nhuge);
            TC_HIP(hipMemcpy(rows.data(), lst + hoff[4], (size_t)nhuge * sizeof(int32_t),
                             hipMemcpyDeviceToHost));
            for (int32_t r : rows) {
                VInfo x;
                TC_HIP(hipMemcpy(&x, vi + r, sizeof(VInfo), hipMemcpyDeviceToHost));
                g.plan.huge.emplace_back(x.rb, (int64_t)(x.len - x.dm));
            }
        }
    }
    if (!g.plan.huge.empty()) {
        // rows longer than any LDS tile: radix sort each one in global memory
        const int vbits = std::max(1, ceil_log2_i((int64_t)nv));
        for (const auto &row : g.plan.huge) {
            const int64_t dp = row.second;
            TC_TRY(g.ov[9].ensure((size_t)dp * sizeof(uint64_t)));
            uint64_t *k64 = g.ov[9].as<uint64_t>();
            int32_t *P = padj + row.first;
            hipLaunchKernelGGL(k_widen_u32, dim3(grid_for(dp, 256, 4096)), dim3(256), 0, s, P, dp, k64);
            TC_TRY(radix_sort_u64(k64, nullptr, dp, vbits, g.ov[4], g.ov[5], g.tmp, s));
            hipLaunchKernelGGL(k_narrow_u64, dim3(grid_for(dp, 256, 4096)), dim3(256), 0, s, k64, dp, P);
            TC_HIP(hipGetLastError());
        }
    }

    // ---- task array: (task vertex, suffix start) for every owner
    const int64_t pchunks = ceil_div(nplus, kWave), tchunks = ceil_div(ntasks, kWave);
    TC_TRY(g.ov[10].ensure((size_t)(pchunks + 1) * sizeof(int32_t)));
    int32_t *pcrow = g.ov[10].as<int32_t>();
    hipLaunchKernelGGL(k_chunk_rows, dim3(vblocks), dim3(256), 0, s, pcoff, nv, pcrow);
    TC_TRY(g.crow.ensure((size_t)(tchunks + 1) * sizeof(int32_t)));
    hipLaunchKernelGGL(k_chunk_rows, dim3(vblocks), dim3(256), 0, s, toff, nv, g.crow.as<int32_t>());
    int32_t *cursor = dmo;  // store-order |N-| is not needed any more
    int32_t *bad = lcnt + 5;
    // the partitioned transpose's bound on its fine buckets (window starts,
    // owners of more than a window and the owners after them) and its
    // staging (16 bytes per task); the appends when either does not fit
    const int64_t fmax = std::min<int64_t>(nv, 3 * (ntasks / kTpWin) + 1) + 1;
    bool part = !fallback && nplus > 0 && fmax <= (int64_t)kTpMaxB * kTpMaxB;
    if (part) {
        const int64_t tmax = ceil_div(ntasks, kTpTile) + kTpMaxB;  // pass-2 tiles, at most
        const size_t meta =
            (size_t)(4 + kTpMaxB + fmax) * 8 + (size_t)(fmax + 1 + kTpMaxB + 1 + tmax) * 4;
        part = g.ov[13].ensure((size_t)(nv + 1) * sizeof(int64_t)) == TC_OK &&
               g.ov[14].ensure(meta) == TC_OK &&
               g.ov[15].ensure((size_t)ntasks * sizeof(uint64_t)) == TC_OK &&
               g.ov[16].ensure((size_t)ntasks * sizeof(int32_t)) == TC_OK &&
               g.ov[17].ensure((size_t)ntasks * sizeof(int32_t)) == TC_OK;
        if (!part)  // out of memory for the staging: free it again
            for (int q = 13; q < 18; ++q) g.ov[q].release();
    }
    // hubs: owners whose task lists are built from their rows (the appends
    // would serialise on their cursors; the partitioned transpose has no
    // per-owner atomics and needs no hubs unless the schedule asks)
    hipLaunchKernelGGL(k_hub_lo, dim3(1), dim3(1), 0, s, vi, nv,
                       hub_degree > 0 ? hub_degree : part ? INT32_MAX : kHubDegree, sc);
    TC_HIP(hipGetLastError());
    if (part) {
        int64_t *fpos = g.ov[13].as<int64_t>();
        int64_t *ctl = g.ov[14].as<int64_t>();
        unsigned long long *ccur = reinterpret_cast<unsigned long long *>(ctl + 4);
        unsigned long long *fcur = ccur + kTpMaxB;
        int32_t *fb = reinterpret_cast<int32_t *>(fcur + fmax);
        int32_t *tiles = fb + fmax + 1;
        int32_t *tmap = tiles + kTpMaxB + 1;
        uint64_t *rec1 = g.ov[15].as<uint64_t>();
        int32_t *own1 = g.ov[16].as<int32_t>(), *own2 = g.ov[17].as<int32_t>();
        uint64_t *tasks = reinterpret_cast<uint64_t *>(g.tasks.as<TaskRec>());
        int32_t *flag = cursor;
        hipLaunchKernelGGL(k_tp_flags, dim3(vblocks), dim3(256), 0, s, toff, sc, nv, flag);
        TC_TRY(scan_counts_i32(flag, fpos, nv, g.tmp, s));
        hipLaunchKernelGGL(k_tp_bounds, dim3(vblocks), dim3(256), 0, s, flag, fpos, sc, nv, fb);
        hipLaunchKernelGGL(k_tp_plan, dim3(1), dim3(1024), 0, s, fpos, nv, fb, toff, ctl, tiles, ccur);
        hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(fmax, 256, 4096)), dim3(256), 0, s, fcur, fmax);
        hipLaunchKernelGGL(k_tp_coarse, dim3((unsigned)ceil_div(nplus, kTpTile)), dim3(kTpThreads), 0, s,
                           pcoff, pcrow, padj, toff, fb, ctl, nv, nplus, sc, ccur, rec1, own1, bad);
        hipLaunchKernelGGL(k_tp_tilemap, dim3(kTpMaxB), dim3(256), 0, s, ctl, tiles, tmap);
        hipLaunchKernelGGL(k_tp_fine, dim3((unsigned)(ceil_div(ntasks, kTpTile) + kTpMaxB)),
                           dim3(kTpThreads), 0, s, rec1, own1, toff, fb, ctl, tiles, tmap, fcur, tasks,
                           own2, bad);
        hipLaunchKernelGGL(k_tp_local, dim3((unsigned)std::min<int64_t>(fmax, 1 << 16)),
                           dim3(kTpThreads), 0, s, toff, fb, ctl, tasks, own2, bad);
        hipLaunchKernelGGL(k_orient_transpose_hubs, dim3(cus * 16), dim3(256), 0, s, pcoff, toff,
                           order, rnk, ocoff, g.st.boff.as<int64_t>(), g.st.pool.as<int32_t>(), mask,
                           cscan, padj, nv, sc, g.tasks.as<TaskRec>(), bad);
        hipLaunchKernelGGL(k_tp_check, dim3(grid_for(fmax, 256, 4096)), dim3(256), 0, s, toff, fb, ctl,
                           ccur, fcur, nv, sc, bad);
    } else {
        hipLaunchKernelGGL(k_fill_u32, dim3(vblocks), dim3(256), 0, s,
                           reinterpret_cast<uint32_t *>(cursor), (int64_t)nv, 0u);
        if (nplus > 0) {
            hipLaunchKernelGGL(k_orient_transpose, dim3(grid_for(pchunks, 4, 8192)), dim3(256), 0, s,
                               pcoff, pcrow, padj, toff, nv, nplus, g.tasks.as<TaskRec>(), cursor,
                               sc, bad);
            hipLaunchKernelGGL(k_orient_transpose_hubs, dim3(cus * 16), dim3(256), 0, s, pcoff, toff,
                               order, rnk, ocoff, g.st.boff.as<int64_t>(), g.st.pool.as<int32_t>(),
                               mask, cscan, padj, nv, sc, g.tasks.as<TaskRec>(), bad);
        }
        hipLaunchKernelGGL(k_check_cursor, dim3(vblocks), dim3(256), 0, s, cursor, vi, toff, nv, sc,
                           bad);
    }
    TC_HIP(hipGetLastError());
    g.plan.bad = bad;
    int32_t hbad = 0;
    if (!g.capturing) {  // a recorded count checks the flag after its replay
        TC_HIP(hipMemcpyAsync(&hbad, bad, sizeof(int32_t), hipMemcpyDeviceToHost, s));
        TC_HIP(hipStreamSynchronize(s));
    }
    if (hbad) {
        g.symmetric = 0;
        set_last_error(TC_ERR_INVALID, "oriented layout: the rows are not symmetric");
        return TC_ERR_INVALID;
    }
    g.ntasks = ntasks;
    g.nplus = nplus;
    g.oriented_views = true;
    return TC_OK;
}

extern const uint8_t *learned_table_device(int32_t *dim);

namespace {

// The per-bin counters of a count (accum words, see count_pass) into stats.
void fill_counters(tc_stats &st, const unsigned long long *hacc, const int64_t *jt, bool oriented)
{
    for (int b = 0; b < 3; ++b) {
        st.jobs_bin[b] = jt[b];
        st.tasks_bin[b] = (int64_t)hacc[4 + 2 * b];
        st.bytes_alg_bin[b] = (double)hacc[5 + 2 * b];
        st.items_bin[b] = (double)hacc[10 + b];
        st.stream_bytes_bin[b] = 4.0 * (double)hacc[10 + b] + 24.0 * (double)hacc[4 + 2 * b] +
                                 4.0 * (double)hacc[13 + b];
    }
    for (int q = 0; q < 3; ++q) st.strategy_tasks[q] = (int64_t)hacc[1 + q];
    st.mergepath_tasks = (int64_t)hacc[25];
    if (oriented && jt[2] > 0) {
        for (int k = 0; k < 2; ++k) {
            const unsigned long long *x = hacc + 16 + 3 * k;
            st.tasks_long[k] = (int64_t)x[0];
            st.items_long[k] = (double)x[1];
            st.stream_bytes_long[k] = 4.0 * (double)x[1] + 24.0 * (double)x[0] + 4.0 * (double)x[2];
        }
    }
    st.triangles = oriented ? hacc[0] / 6 : 0;
}

// One count.  While g.capturing (stream capture of a hipGraph), the values
// the host would read back come from g.plan, and no event is recorded and
// nothing is read back: the caller reads the results after the replay.
int count_pass(Graph &g, uint32_t *out, uint64_t *global, const tc_schedule &sched,
               tc_stats *stats, bool skip_lower)
{
    const int32_t nv = g.st.nv;
    const int64_t nnz = g.st.nnz;
    hipStream_t s = g.stream;
    TC_TRY(g.ensure_events());
    tc_stats st = {};
    auto rec = [&](hipEvent_t e) { return g.capturing ? hipSuccess : hipEventRecord(e, s); };

    int method = sched.method;
    if (skip_lower) method = TC_METHOD_DIRECT;
    if (method == TC_METHOD_AUTO || method == TC_METHOD_ORIENTED) {
        int64_t bad = 0;
        if (g.symmetric < 0 || g.self_loops < 0) TC_TRY(validate_graph(g, true, &bad));
        bool ok = g.symmetric == 1 && g.self_loops == 0;
        if (method == TC_METHOD_ORIENTED && !ok) {
            set_last_error(TC_ERR_INVALID,
                           "oriented method needs a symmetric graph without self-loops");
            return TC_ERR_INVALID;
        }
        method = ok ? TC_METHOD_ORIENTED : TC_METHOD_DIRECT;
    }
    bool oriented = method == TC_METHOD_ORIENTED;

    // scratch: accum[0] = global, [1..3] = strategy counts, [4..9] bin tasks
    // and bytes, [10..12] items, [13..15] tile items, [16..21] long-task stats,
    // [22..24] job counters of the largest bin's two launches, [25] merge-path
    // tasks
    TC_TRY(g.accum.ensure(32 * sizeof(unsigned long long)));
    unsigned long long *acc = g.accum.as<unsigned long long>();
    uint32_t *outp = out;
    if (!outp) {
        TC_TRY(g.tmp4.ensure((size_t)std::max<int32_t>(nv, 1) * sizeof(uint32_t)));
        outp = g.tmp4.as<uint32_t>();
    }

    TC_HIP(rec(g.ev[0]));
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, s, acc, (int64_t)32);
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(nv, 256, 65536)), dim3(256), 0, s, outp,
                       (int64_t)nv, 0u);
    TC_HIP(hipGetLastError());

    if (nv == 0 || nnz == 0) {
        TC_HIP(hipEventRecord(g.ev[1], s));
        TC_HIP(hipEventSynchronize(g.ev[1]));
        if (global) *global = 0;
        st.method_used = method;
        if (stats) *stats = st;
        return TC_OK;
    }

    // compact offsets, chunk rows and the (degree-partitioned) vertex views;
    // rows declared symmetric that turn out not to be go the direct way
    if (oriented) {
        const int pv = prepare_views(g, true, sched.hub_degree, sched.build);
        if (pv == TC_ERR_INVALID && sched.method == TC_METHOD_AUTO && !g.symmetric) {
            oriented = false;
            method = TC_METHOD_DIRECT;
        } else if (pv != TC_OK) {
            return pv;
        }
    }
    if (!oriented) TC_TRY(prepare_views(g, false));
    st.method_used = method;
    int64_t *coff = g.coff.as<int64_t>();
    int32_t *crow = g.crow.as<int32_t>();
    VInfo *vi = g.vinfo.as<VInfo>();
    const int32_t *adj = oriented ? g.padj.as<int32_t>() : g.st.pool.as<int32_t>();
    const TaskRec *tasks = oriented ? g.tasks.as<TaskRec>() : nullptr;
    // planning walks the task array (oriented) or the store rows (direct)
    const int64_t nent = oriented ? g.ntasks : nnz;
    const int64_t nchunks = ceil_div(nent, kWave);
    // 4 waves of 256 threads; 16 blocks per CU keep the chip busy and bound
    // the number of per-block statistics atomics
    const int wblocks = grid_for(nchunks, 4, 4096);

    // ---- planning: work per owner, jobs per bin
    PlanParams P;
    P.lmax[0] = kL0;
    P.lmax[1] = kL1;
    P.lmax[2] = kL2;
    P.nw[0] = kNW0;
    P.nw[1] = kNW1;
    P.nw[2] = kNW2;
    P.bin_max[0] = std::min(std::max(sched.small_max, 1), kL0);
    P.bin_max[1] = std::min(std::max(sched.mid_max, P.bin_max[0]), kL1);
    P.job_items = std::max(sched.job_items, 256);
    P.flush_items = sched.flush_items > 0 ? sched.flush_items : 1024;
    TC_TRY(g.work.ensure((size_t)nv * sizeof(unsigned long long)));
    unsigned long long *work = g.work.as<unsigned long long>();
    hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(nv, 256, 65536)), dim3(256), 0, s, work, (int64_t)nv);
    if (oriented)
        hipLaunchKernelGGL(k_owner_work<true>, dim3(wblocks), dim3(256), 0, s, coff, crow, vi, adj,
                           tasks, nv, nent, P, work, stats ? acc + 4 : nullptr);
    else
        hipLaunchKernelGGL(k_owner_work<false>, dim3(wblocks), dim3(256), 0, s, coff, crow, vi, adj,
                           nullptr, nv, nent, P, work, stats ? acc + 4 : nullptr);
    TC_HIP(hipGetLastError());
    TC_TRY(g.tmp3.ensure((size_t)3 * nv * sizeof(int32_t) + (size_t)3 * (nv + 1) * sizeof(int64_t) + 16));
    int32_t *njobs = g.tmp3.as<int32_t>();
    int64_t *jscan = reinterpret_cast<int64_t *>(njobs + 3 * (int64_t)nv + ((3 * (int64_t)nv) & 1));
    if (oriented)
        hipLaunchKernelGGL(k_plan_jobs<true>, dim3(grid_for(nv, 256, 4096)), dim3(256), 0, s, vi,
                           work, nv, P, njobs, stats ? acc + 13 : nullptr);
    else
        hipLaunchKernelGGL(k_plan_jobs<false>, dim3(grid_for(nv, 256, 4096)), dim3(256), 0, s, vi,
                           work, nv, P, njobs, stats ? acc + 13 : nullptr);
    TC_HIP(hipGetLastError());
    for (int b = 0; b < 3; ++b)
        TC_TRY(scan_counts_i32(njobs + (int64_t)b * nv, jscan + (int64_t)b * (nv + 1), nv, g.tmp, s));
    int64_t jt[3] = {g.plan.jt[0], g.plan.jt[1], g.plan.jt[2]};
    if (!g.capturing) {
        for (int b = 0; b < 3; ++b)
            TC_HIP(hipMemcpyAsync(&jt[b], jscan + (int64_t)b * (nv + 1) + nv, sizeof(int64_t),
                                  hipMemcpyDeviceToHost, s));
        TC_HIP(hipStreamSynchronize(s));
        for (int b = 0; b < 3; ++b) g.plan.jt[b] = jt[b];
    }
    TC_TRY(g.jobs.ensure((size_t)(jt[0] + jt[1] + jt[2] + 1) * sizeof(Job)));
    Job *j0 = g.jobs.as<Job>();
    Job *j1 = j0 + jt[0];
    Job *j2 = j1 + jt[1];
    if (oriented)
        hipLaunchKernelGGL(k_emit_jobs<true>, dim3(grid_for(nv, 256, 65536)), dim3(256), 0, s, vi,
                           njobs, jscan, nv, P, j0, j1, j2);
    else
        hipLaunchKernelGGL(k_emit_jobs<false>, dim3(grid_for(nv, 256, 65536)), dim3(256), 0, s, vi,
                           njobs, jscan, nv, P, j0, j1, j2);
    TC_HIP(hipGetLastError());
    TC_HIP(rec(g.ev[2]));

    // ---- count
    CountArgs A = {};
    A.adj = adj;
    A.vi = vi;
    A.tasks = tasks;
    A.pcnt = nullptr;
    A.toff = oriented ? coff : nullptr;
    A.order = oriented ? g.order.as<int32_t>() : nullptr;
    A.out = outp;
    A.global = acc;
    A.strat = acc + 1;
    A.strat_mp = acc + 25;
    A.long_stats = acc + 16;
    A.job_ctr = acc + 22;
    A.policy = sched.policy;
    A.ratio_x16 = std::max(sched.ratio_x16, 0);
    A.mpath_min = std::max(sched.mpath_min, 0);
    A.skip_lower = skip_lower ? 1 : 0;
    A.per_graph = g.batch_per_graph;
    A.gid = nullptr;
    if (g.batch_per_graph) {
        // graph of every owner, found once here instead of once per job
        TC_TRY(g.ov[11].ensure((size_t)std::max(nv, 1) * sizeof(int32_t)));
        hipLaunchKernelGGL(k_graph_of, dim3(grid_for(nv, 256, 65536)), dim3(256), 0, s, g.batch_voff,
                           g.batch_ngraphs, oriented ? g.order.as<int32_t>() : nullptr, nv,
                           g.ov[11].as<int32_t>());
        TC_HIP(hipGetLastError());
        A.gid = g.ov[11].as<int32_t>();
    }
    A.table = nullptr;
    A.table_dim = 0;
    if (sched.policy == TC_POLICY_LEARNED) {
        int32_t dim = 0;
        A.table = learned_table_device(&dim);
        A.table_dim = dim;
    }
    const int64_t nplus = oriented ? g.nplus : 0;
    if (oriented) {
        TC_TRY(g.pcnt.ensure((size_t)std::max<int64_t>(nplus, 1) * sizeof(uint32_t)));
        hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(nplus, 256, 65536)), dim3(256), 0, s,
                           g.pcnt.as<uint32_t>(), nplus, 0u);
        A.pcnt = g.pcnt.as<uint32_t>();
        TC_TRY(g.outr.ensure((size_t)std::max<int32_t>(nv, 1) * sizeof(uint32_t)));
        hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(nv, 256, 65536)), dim3(256), 0, s,
                           g.outr.as<uint32_t>(), (int64_t)nv, 0u);
        A.outr = g.outr.as<uint32_t>();
    }
    if (oriented && jt[2] > 0) {
        TC_TRY(g.hjobs.ensure((size_t)jt[2] * sizeof(int64_t)));
        A.hjobs = g.hjobs.as<int64_t>();
    }
    const Job *jl[3] = {j0, j1, j2};
    for (int b = 0; b < 3; ++b) {
        TC_HIP(rec(g.ev[3 + b]));
        TC_TRY(launch_bin(b, oriented, jl[b], jt[b], A, s, g.capturing ? nullptr : &g.ev[8]));
    }
    TC_HIP(rec(g.ev[6]));  // end of the bin kernels
    if (oriented && nplus > 0) {
        hipLaunchKernelGGL(k_plus_row_sum, dim3(grid_for(ceil_div(nplus, kWave), 4, 1 << 16)),
                           dim3(256), 0, s, g.ov[6].as<int64_t>(), g.ov[10].as<int32_t>(),
                           g.pcnt.as<uint32_t>(), nv, nplus, g.outr.as<uint32_t>());
        TC_HIP(hipGetLastError());
    }
    if (oriented) {
        hipLaunchKernelGGL(k_rank_to_vertex, dim3(grid_for(nv, 256, 65536)), dim3(256), 0, s,
                           g.order.as<int32_t>(), nv, g.outr.as<uint32_t>(), outp);
        TC_HIP(hipGetLastError());
    }
    g.plan.method = method;
    if (g.capturing) return TC_OK;  // results are read after the replay
    TC_HIP(hipEventRecord(g.ev[7], s));
    TC_HIP(hipEventSynchronize(g.ev[7]));

    unsigned long long hacc[26];
    TC_HIP(hipMemcpy(hacc, acc, sizeof(hacc), hipMemcpyDeviceToHost));
    if (global) *global = hacc[0];
    if (stats) {
        fill_counters(st, hacc, jt, oriented);
        float ms = 0;
        (void)hipEventElapsedTime(&ms, g.ev[0], g.ev[7]);
        st.ms_total = ms;
        (void)hipEventElapsedTime(&ms, g.ev[0], g.ev[2]);
        st.ms_prepare = ms;
        (void)hipEventElapsedTime(&ms, g.ev[3], g.ev[7]);
        st.ms_count = ms;
        for (int b = 0; b < 3; ++b) {
            (void)hipEventElapsedTime(&ms, g.ev[3 + b], g.ev[4 + b]);
            st.ms_bin[b] = ms;
        }
        if (oriented && jt[2] > 0) {
            for (int k = 0; k < 2; ++k) {
                (void)hipEventElapsedTime(&ms, g.ev[8 + k], g.ev[9 + k]);
                st.ms_long[k] = ms;
            }
        }
        *stats = st;
    }
    return TC_OK;
}

// Key of a recorded count: everything its launches bake in (schedule,
// output, learned table, statistics on or off, the store's version).
uint64_t plan_key(const Graph &g, const uint32_t *out, const tc_schedule &sched, bool with_stats,
                  bool skip_lower)
{
    tc_schedule k = sched;
    k.capture = 0;
    int32_t dim = 0;
    const uint8_t *table = sched.policy == TC_POLICY_LEARNED ? learned_table_device(&dim) : nullptr;
    uint64_t h = 1469598103934665603ull;
    auto mix = [&](const void *p, size_t n) {
        const unsigned char *b = static_cast<const unsigned char *>(p);
        for (size_t i = 0; i < n; ++i) h = (h ^ b[i]) * 1099511628211ull;
    };
    mix(&k, sizeof(k));
    mix(&out, sizeof(out));
    mix(&table, sizeof(table));
    mix(&g.version, sizeof(g.version));
    mix(&g.batch_per_graph, sizeof(g.batch_per_graph));  // batched handles: per-graph output
    mix(&g.batch_voff, sizeof(g.batch_voff));
    const int32_t flags = (with_stats ? 1 : 0) | (skip_lower ? 2 : 0);
    mix(&flags, sizeof(flags));
    return h | 1ull;  // never 0 (no recorded graph)
}

// Launches the recorded count and reads its results.
int count_replay(Graph &g, uint64_t *global, tc_stats *stats)
{
    hipStream_t s = g.stream;
    TC_HIP(hipEventRecord(g.ev[0], s));
    TC_HIP(hipGraphLaunch(g.cexec, s));
    TC_HIP(hipEventRecord(g.ev[7], s));
    TC_HIP(hipEventSynchronize(g.ev[7]));
    unsigned long long hacc[26];
    int32_t bad = 0;
    TC_HIP(hipMemcpy(hacc, g.accum.p, sizeof(hacc), hipMemcpyDeviceToHost));
    if (g.plan.method == TC_METHOD_ORIENTED && g.plan.bad)
        TC_HIP(hipMemcpy(&bad, g.plan.bad, sizeof(bad), hipMemcpyDeviceToHost));
    if (bad) {  // cannot happen for an unchanged store that counted before
        g.symmetric = 0;
        g.touch();
        set_last_error(TC_ERR_INVALID, "oriented layout: the rows are not symmetric");
        return TC_ERR_INVALID;
    }
    if (global) *global = hacc[0];
    if (stats) {
        tc_stats st = {};
        fill_counters(st, hacc, g.plan.jt, g.plan.method == TC_METHOD_ORIENTED);
        float ms = 0;
        (void)hipEventElapsedTime(&ms, g.ev[0], g.ev[7]);
        st.ms_total = ms;
        st.method_used = g.plan.method;
        st.graph_replay = 1;
        *stats = st;
    }
    return TC_OK;
}

// Records one count as a hipGraph on a stream of its own (the legacy default
// stream cannot be captured), using the plan of the count before.
int count_capture(Graph &g, uint32_t *out, const tc_schedule &sched, tc_stats *stats, bool skip_lower,
                  uint64_t key)
{
    if (!g.cap_stream) TC_HIP(hipStreamCreateWithFlags(&g.cap_stream, hipStreamNonBlocking));
    TC_HIP(hipStreamSynchronize(g.stream));
    hipStream_t user = g.stream;
    g.stream = g.cap_stream;
    g.capturing = true;
    hipGraph_t graph = nullptr;
    int st = TC_OK;
    hipError_t e = hipStreamBeginCapture(g.cap_stream, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        tc_stats dummy;
        st = count_pass(g, out, nullptr, sched, stats ? &dummy : nullptr, skip_lower);
        e = hipStreamEndCapture(g.cap_stream, &graph);
    }
    g.capturing = false;
    g.stream = user;
    if (st == TC_OK && e != hipSuccess) {
        set_last_error(TC_ERR_HIP, std::string("count capture: ") + hipGetErrorString(e));
        st = TC_ERR_HIP;
    }
    if (st != TC_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        g.plan.valid = false;
        return st;
    }
    hipGraphExec_t exec = nullptr;
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(graph);
        g.plan.valid = false;
        set_last_error(TC_ERR_HIP, std::string("count graph: ") + hipGetErrorString(e));
        return TC_ERR_HIP;
    }
    g.drop_capture();
    g.cgraph = graph;
    g.cexec = exec;
    g.cexec_key = key;
    return TC_OK;
}

}  // namespace

int static_count(Graph &g, uint32_t *out, uint64_t *global, const tc_schedule &sched,
                 tc_stats *stats, bool skip_lower)
{
    if (!sched.capture || g.st.nv == 0 || g.st.nnz == 0)
        return count_pass(g, out, global, sched, stats, skip_lower);
    const uint64_t key = plan_key(g, out, sched, stats != nullptr, skip_lower);
    if (g.cexec && g.cexec_key == key) return count_replay(g, global, stats);
    if (g.plan.valid && g.plan.key == key) {
        if (count_capture(g, out, sched, stats, skip_lower, key) == TC_OK)
            return count_replay(g, global, stats);
        (void)hipGetLastError();  // not recordable here: count as usual
        return count_pass(g, out, global, sched, stats, skip_lower);
    }
    // first count with this key: a plain pass that leaves its plan behind
    g.drop_capture();
    g.plan.valid = false;
    const int st = count_pass(g, out, global, sched, stats, skip_lower);
    if (st == TC_OK) {
        g.plan.valid = true;
        g.plan.key = key;
    }
    return st;
}

}  // namespace tc
