// tc_internal.hpp -- shared host/device definitions of the triangle-counting
// library (not part of the public C-ABI; see include/custinger_tc.h).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/custinger_tc.h"

namespace tc {

constexpr int kWave = 64;

// ---------------------------------------------------------------- errors --
struct Error {
    int code;
    std::string msg;
};

void set_last_error(int code, const std::string &msg);

#define TC_HIP(call)                                                                   \
    do {                                                                               \
        hipError_t _e = (call);                                                        \
        if (_e != hipSuccess) {                                                        \
            ::tc::set_last_error(TC_ERR_HIP, std::string(#call) + ": " +               \
                                                 hipGetErrorString(_e));               \
            return TC_ERR_HIP;                                                         \
        }                                                                              \
    } while (0)

#define TC_TRY(expr)                                                                   \
    do {                                                                               \
        int _s = (expr);                                                               \
        if (_s != TC_OK) return _s;                                                    \
    } while (0)

// ------------------------------------------------------- device buffers --
// Scratch allocations are owned by the graph handle and grown on demand.
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    int ensure(size_t need);
    void release();
    template <class T> T *as() const { return static_cast<T *>(p); }
};

// ------------------------------------------------------------- helpers --
__host__ __device__ inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

inline int grid_for(int64_t n, int block, int64_t cap = 1 << 20)
{
    int64_t g = ceil_div(n, block);
    if (g < 1) g = 1;
    if (g > cap) g = (int)cap;
    return (int)g;
}

// Device-wide exclusive scan of int64 values (in may equal out).  total, if
// non-null, receives the sum on the device (one int64).  Uses tmp scratch.
int scan_exclusive_i64(const int64_t *in, int64_t *out, int64_t n, int64_t *total,
                       DevBuf &tmp, hipStream_t s);
// int32 counts -> int64 exclusive prefix (out has n+1 entries, out[n]=sum).
int scan_counts_i32(const int32_t *in, int64_t *out, int64_t n, DevBuf &tmp, hipStream_t s);

// LSD radix sort of uint64 keys over bits [0, bits).  Keys and an optional
// int32 payload are sorted in place (alt buffers supplied by scratch).
int radix_sort_u64(uint64_t *keys, int32_t *vals, int64_t n, int bits, DevBuf &alt_keys,
                   DevBuf &alt_vals, DevBuf &tmp, hipStream_t s);

// ------------------------------------------------------- vertex views --
// One 16-byte record per vertex used by every intersection kernel.  Direct
// layout: the row of v is adj[rb .. rb+len) and dm = 0.  Oriented (rank)
// layout: len is the degree, dm = |N-(v)| (the tasks of v, kept in a separate
// task array) and N+(v) = adj[rb .. rb+len-dm), sorted by rank.
struct VInfo {
    int64_t rb;
    int32_t dm;
    int32_t len;
};

// One task of the oriented layout: for task vertex a of an owner b, the
// suffix of N+(a) ranked above b -- the only entries that can close a
// triangle -- as its start in the plus array (bits 24..63) and its length
// (bits 0..23).  Under the degree order |N+(a)| <= sqrt(nnz), far below 2^24,
// and the plus array has fewer than 2^40 entries.  a itself is the plus-array
// row that holds entry base - 1.
struct TaskRec {
    uint64_t v;

    __host__ __device__ static TaskRec make(int64_t base, int32_t len)
    {
        return TaskRec{((uint64_t)base << 24) | (uint64_t)(uint32_t)len};
    }
    __host__ __device__ int64_t base() const { return (int64_t)(v >> 24); }
    __host__ __device__ int32_t len() const { return (int32_t)(v & 0xFFFFFFu); }
};

// Intersection jobs: a slice [t_begin, t_end) of an owner's task list and the
// tile of the owner's list that sits in LDS.
struct Job {
    int32_t owner;
    int32_t t_begin;
    int32_t t_end;
    int32_t tile;
};

// ------------------------------------------------------------- store --
// Dynamic adjacency store: vertex v owns a block pool[boff[v] .. boff[v]+cap[v])
// of which the first len[v] entries are used, sorted ascending.
struct Store {
    int32_t nv = 0;
    int64_t nnz = 0;           // sum of len
    int64_t pool_used = 0;     // high-water mark of allocated blocks
    int64_t reallocs = 0;
    DevBuf boff, len, cap, pool;
};

struct Graph {
    Store st;
    hipStream_t stream = nullptr;
    bool validated = false;
    int symmetric = -1;        // -1 unknown
    int64_t self_loops = -1;   // -1 unknown
    int32_t max_degree = 0;
    // cached derived structures (rebuilt after updates)
    DevBuf vinfo, padj, coff, crow, work, jobs, tmp, tmp2, tmp3, tmp4, accum;
    // oriented layout in rank space: order[r] = vertex of rank r, rnk = its
    // inverse; tasks = for every owner b (rows by rank, offsets in coff) the
    // pairs (a, position of b in the rank-sorted N+(a)) for a in N-(b);
    // ov = scratch of the orientation
    DevBuf order, rnk, tasks;
    DevBuf hjobs;  // the largest bin's jobs left to its hash-table launch
    DevBuf outr;   // oriented per-vertex counts by rank, moved to vertex order after the count
    DevBuf ov[19];
    int64_t ntasks = 0;  // entries of the task array (oriented layout)
    int64_t nplus = 0;   // entries of the plus array (oriented layout)
    DevBuf pcnt;         // per N+ entry: hits of the task it became (summed per vertex)
    bool oriented_views = false;
    DevBuf bsrc, bdst, bflag, bkeys, bkeys2, bvals2, boffb, bdel;
    // streaming: per-entry "in the current batch" flags parallel to the pool,
    // per-vertex marks, and batch scratch
    DevBuf flagpool, vmark;
    DevBuf sb[28];
    // capacity of a row that outgrows its block: len * grow_mul_x16 / 16 +
    // grow_add (the update allocator's policy); 0: the next power of two
    int32_t grow_mul_x16 = 0;
    int32_t grow_add = 0;
    DevBuf lbuf[5];  // delta: items, their prefix, per-edge counts, merge segments, their prefix
    bool degree_stale = false;
    // batched counting: graph boundaries and per-graph output (device)
    const int64_t *batch_voff = nullptr;
    int32_t batch_ngraphs = 0;
    unsigned long long *batch_per_graph = nullptr;
    // the pool is the caller's adjacency array (read-only, not freed)
    bool borrowed = false;
    hipEvent_t ev[16] = {};
    bool events_ready = false;
    int ensure_events();
    // pinned host words the batch sequence's counts are copied to (32)
    unsigned long long *hcounts = nullptr;
    // store version: bumped by every update, so that plans and recorded
    // counts of an older store are never reused
    uint64_t version = 0;
    // hipGraph replay of static counts (tc_schedule.capture).  A count
    // records here the values the host reads back mid-sequence; a count with
    // the same key then runs the sequence under stream capture with these
    // values (capturing = true: no host read-backs, no events), and later
    // ones launch the instantiated graph.
    struct CountPlan {
        bool valid = false;
        uint64_t key = 0;
        int method = 0;
        unsigned long long hmax = 0;        // maximum degree (rank-sort bits)
        int64_t nplus = 0, ntasks = 0;      // plus-array and task-array entries
        std::vector<std::pair<int64_t, int64_t>> huge;  // N+ rows sorted in global memory: (rb, len)
        int64_t jt[3] = {0, 0, 0};          // jobs per bin
        int32_t *bad = nullptr;             // the orientation's symmetry flag (checked after a replay)
    } plan;
    bool capturing = false;
    hipStream_t cap_stream = nullptr;
    hipGraph_t cgraph = nullptr;
    hipGraphExec_t cexec = nullptr;
    uint64_t cexec_key = 0;
    // recorded batch sequences (tc_batch_apply with tc_schedule.capture):
    // one per size class and buffer layout, keyed like the counts
    hipGraph_t bgraph = nullptr;
    hipGraphExec_t bexec = nullptr;
    uint64_t bexec_key = 0, bplan_key = 0;
    bool bplan_valid = false;
    bool bcap_failed = false;  // a recording failed: batches run plainly
    void drop_capture();
    void touch()  // the store changed
    {
        ++version;
        plan.valid = false;
        drop_capture();
    }
};

// Builds coff, crow and vinfo (and padj for the oriented layout) for the store;
// hub_degree > 0 overrides the degree from which owners build their task
// lists from their own rows; fallback = 1 takes the paths used when the
// default's staging does not fit (per-entry appends, mask by degree).
int prepare_views(Graph &g, bool oriented, int32_t hub_degree = 0, int fallback = 0);
// Fills vinfo with the store's rows as given (direct layout, dm=0).
int build_direct_vinfo(Graph &g);
// Validation of sortedness/range, self-loop count, optional symmetry check.
int validate_graph(Graph &g, bool check_symmetry, int64_t *bad_rows);

int static_count(Graph &g, uint32_t *out, uint64_t *global, const tc_schedule &sched,
                 tc_stats *stats, bool skip_lower = false);
int graph_create_csr(int32_t nv, int64_t nnz, const int64_t *off, const int32_t *adj, int location,
                     uint32_t flags, int32_t slack_pct, Graph **out);
int graph_to_csr(Graph &g, int64_t *off, int32_t *adj, int location);
int graph_borrow_csr(int32_t nv, int64_t nnz, const int64_t *d_off, const int32_t *d_adj,
                     hipStream_t s, Graph **out);
void graph_destroy(Graph *g);
tc_schedule default_schedule();

// ------------------------------------------------------- reductions --
template <class T> __device__ inline T wave_sum(T x)
{
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) x += __shfl_xor(x, d, kWave);
    return x;
}

// Adds the block's sums of x[0..N) to dst[0..N) with one atomic per nonzero
// sum per block (device-scope atomics on one address serialise, so a grid of
// many waves must not issue one each).  Every thread of the block calls it;
// blockDim.x is a multiple of the wave size, at most 1024.
template <int N>
__device__ inline void block_add_u64(const unsigned long long (&x)[N], unsigned long long *dst)
{
    __shared__ unsigned long long part[16][N];
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
#pragma unroll
    for (int q = 0; q < N; ++q) {
        const unsigned long long t = wave_sum<unsigned long long>(x[q]);
        if (lane == 0) part[wid][q] = t;
    }
    __syncthreads();
    if (threadIdx.x < N) {
        unsigned long long t = 0;
        for (int w = 0; w < (int)(blockDim.x / kWave); ++w) t += part[w][threadIdx.x];
        if (t) atomicAdd(&dst[threadIdx.x], t);
    }
}

// Hash used by the LDS tables (multiplicative, table size 2^bits).
__host__ __device__ inline uint32_t lds_hash(int32_t w, int bits)
{
    return ((uint32_t)w * 0x9E3779B1u) >> (32 - bits);
}

// ---------------------------------------- one-block register bitonic --
// A block of 1024 threads sorts kRegSort keys ascending: thread (wave w,
// lane l) holds positions 128w + l (v0) and 128w + 64 + l (v1).  Partners
// closer than 64 are exchanged by shuffles, those 64 apart are a lane's own
// pair, and only the 10 steps with partners 128 or more apart go through LDS
// (buf, double-buffered: one barrier each).  Every thread calls it.
constexpr int kRegSort = 2048;

__device__ inline uint64_t shfl_xor_u64(uint64_t x, int j)
{
    const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)x, j, 64);
    const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(x >> 32), j, 64);
    return ((uint64_t)hi << 32) | lo;
}

// The value position i keeps at step (k, j) of the network, its partner p.
__device__ inline uint64_t bitonic_keep(int i, int j, int k, uint64_t v, uint64_t p)
{
    const bool asc = (i & k) == 0, lower = (i & j) == 0;
    return lower == asc ? (v < p ? v : p) : (v < p ? p : v);
}

__device__ inline void block_bitonic_2048(uint64_t &v0, uint64_t &v1, uint64_t (*buf)[kRegSort])
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int i0 = 128 * w + lane, i1 = i0 + 64;
    int par = 0;
    for (int k = 2; k <= kRegSort; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            if (j >= 128) {
                buf[par][i0] = v0;
                buf[par][i1] = v1;
                __syncthreads();
                const uint64_t p0 = buf[par][i0 ^ j], p1 = buf[par][i1 ^ j];
                v0 = bitonic_keep(i0, j, k, v0, p0);
                v1 = bitonic_keep(i1, j, k, v1, p1);
                par ^= 1;
            } else if (j == 64) {
                const uint64_t lo = v0 < v1 ? v0 : v1, hi = v0 < v1 ? v1 : v0;
                const bool asc = (i0 & k) == 0;
                v0 = asc ? lo : hi;
                v1 = asc ? hi : lo;
            } else {
                const uint64_t p0 = shfl_xor_u64(v0, j), p1 = shfl_xor_u64(v1, j);
                v0 = bitonic_keep(i0, j, k, v0, p0);
                v1 = bitonic_keep(i1, j, k, v1, p1);
            }
        }
    }
    __syncthreads();  // buf may be reused by the caller
}

}  // namespace tc
