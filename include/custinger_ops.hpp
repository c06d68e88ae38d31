// custinger_ops.hpp -- traversal operators over the dynamic store for user
// HIP kernels: the reference's operator layer (include/operators.cuh,
// include/algs.cuh:14-131, include/load_balance.cuh:35-129) on the C-ABI's
// device view (tc_graph_get_view) and edge numbering (tc_lb_edge_prefix).
//
// Header-only; include it from a .hip source built with
// `hipcc --offload-arch=gfx950` and link libcustinger_tc.so.  Two forms:
//
//  * functors (inlined): for_vertices, for_edges_lb, for_batch_edges, with
//    an operator object called as op(view, v) or op(view, src, dst);
//  * the reference's shape, a __device__ function pointer as template
//    argument and a void* metadata (cusSubKernelVertex / cusSubKernelEdge,
//    algs.cuh:14-16): allVinG_TraverseVertices, allVinA_TraverseVertices,
//    allEinG_TraverseEdges, allVinA_TraverseEdges(_LB), allEinA_TraverseEdges.
//    The functions receive the view (not a device graph pointer).
//
// Edge traversals are load balanced over edges, not vertices: the edges of
// the selected vertices are numbered by a prefix sum of their degrees and
// every wavefront takes 64 consecutive edge numbers, so a hub's row is spread
// over the whole GPU and a run of low-degree vertices shares one wavefront.
// This replaces the reference's 256-vertex block sort into degree bins
// (operators.cuh:165-288).  Every launch goes on the handle's stream.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "custinger_tc.h"

namespace cuStingerAlgs {

// Device view of the store (tc_graph_view without the stream).
struct GraphView {
    int32_t nv;
    int64_t nnz;
    const int64_t *boff;
    const int32_t *len;
    const int32_t *pool;

    __host__ __device__ int32_t degree(int32_t v) const { return len[v]; }
    __host__ __device__ const int32_t *adj(int32_t v) const { return pool + boff[v]; }
};

typedef void (*cusSubKernelVertex)(const GraphView &g, int32_t v, void *metadata);
typedef void (*cusSubKernelEdge)(const GraphView &g, int32_t src, int32_t dst, void *metadata);

namespace ops_detail {

constexpr int kWave = 64;
constexpr int kThreads = 256;

inline int view_of(tc_graph *g, GraphView &v, hipStream_t &s)
{
    tc_graph_view tv;
    const int st = tc_graph_get_view(g, &tv);
    if (st != TC_OK) return st;
    v = GraphView{tv.nv, tv.nnz, tv.boff, tv.len, tv.pool};
    s = (hipStream_t)tv.stream;
    return TC_OK;
}

// Persistent grid: enough blocks to fill the device, never more than the work.
inline int grid(int64_t items, int per_block)
{
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess)
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    int64_t b = (items + per_block - 1) / per_block;
    const int64_t cap = (int64_t)cus * 8;
    if (b > cap) b = cap;
    if (b < 1) b = 1;
    return (int)b;
}

template <class Op>
__global__ __launch_bounds__(kThreads) void k_vertices(GraphView g, const int32_t *verts, int64_t n,
                                                       Op op)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        op(g, verts ? verts[i] : (int32_t)i);
}

// Index t of the vertex whose edges hold edge number x: the last t in [lo, hi]
// with pre[t] <= x.
__device__ inline int64_t owner_of(const int64_t *pre, int64_t x, int64_t lo, int64_t hi)
{
    while (lo < hi) {
        const int64_t mid = lo + ((hi - lo + 1) >> 1);
        if (pre[mid] <= x) lo = mid; else hi = mid - 1;
    }
    return lo;
}

template <class Op>
__global__ __launch_bounds__(kThreads) void k_edges_lb(GraphView g, const int32_t *verts, int64_t n,
                                                       const int64_t *pre, Op op)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t total = pre[n];
    const int64_t nchunks = (total + kWave - 1) / kWave;
    const int64_t wstride = (int64_t)gridDim.x * (blockDim.x / kWave);
    for (int64_t c = (int64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; c < nchunks;
         c += wstride) {
        const int64_t x0 = c * kWave;
        const int64_t xl = (x0 + kWave < total ? x0 + kWave : total) - 1;
        // the chunk's first and last owners bound every lane's search
        int64_t t = 0;
        if (lane == 0) t = owner_of(pre, x0, 0, n - 1);
        if (lane == 1) t = owner_of(pre, xl, 0, n - 1);
        const int64_t tf = __shfl(t, 0, kWave), tl = __shfl(t, 1, kWave);
        const int64_t x = x0 + lane;
        if (x > xl) continue;
        const int64_t ti = tf == tl ? tf : owner_of(pre, x, tf, tl);
        const int32_t src = verts ? verts[ti] : (int32_t)ti;
        const int32_t dst = g.pool[g.boff[src] + (x - pre[ti])];
        op(g, src, dst);
    }
}

template <class Op>
__global__ __launch_bounds__(kThreads) void k_batch_edges(GraphView g, const int32_t *src,
                                                          const int32_t *dst, int64_t n, Op op)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        op(g, src[i], dst[i]);
}

template <cusSubKernelVertex F> struct VertexFn {
    void *meta;
    __device__ void operator()(const GraphView &g, int32_t v) const { F(g, v, meta); }
};
template <cusSubKernelEdge F> struct EdgeFn {
    void *meta;
    __device__ void operator()(const GraphView &g, int32_t s, int32_t d) const { F(g, s, d, meta); }
};

}  // namespace ops_detail

// Scratch of the load-balanced traversals: the edge numbering of the last
// vertex array (cusLoadBalance, include/load_balance.cuh:35-129, keeps its
// work prefix the same way).  One per concurrent traversal.
class LoadBalance {
public:
    LoadBalance() = default;
    LoadBalance(const LoadBalance &) = delete;
    LoadBalance &operator=(const LoadBalance &) = delete;
    ~LoadBalance() { Release(); }

    // Numbers the edges of verts (device ids; null = every vertex).
    int Prepare(tc_graph *g, const int32_t *verts, int64_t n)
    {
        if (!verts) {
            tc_graph_view view;
            const int st = tc_graph_get_view(g, &view);
            if (st != TC_OK) return st;
            n = view.nv;
        }
        if (n + 1 > cap_) {
            Release();
            if (hipMalloc(&pre_, (size_t)(n + 1) * sizeof(int64_t)) != hipSuccess) {
                pre_ = nullptr;
                return TC_ERR_OOM;
            }
            cap_ = n + 1;
        }
        verts_ = verts;
        n_ = n;
        return tc_lb_edge_prefix(g, verts, n, pre_);
    }
    void Release()
    {
        if (pre_) (void)hipFree(pre_);
        pre_ = nullptr;
        cap_ = 0;
    }
    const int64_t *prefix() const { return pre_; }
    const int32_t *vertices() const { return verts_; }
    int64_t size() const { return n_; }

private:
    int64_t *pre_ = nullptr;
    int64_t cap_ = 0;
    const int32_t *verts_ = nullptr;
    int64_t n_ = 0;
};

// Device queue of vertex ids (algs.cuh:40-131's Queue<vertexId_t> on the
// device side): enqueue appends with one atomic per wavefront.
struct DeviceVertexQueue {
    int32_t *items;
    unsigned long long *end;

    __device__ void enqueue(int32_t v) const
    {
        const uint64_t active = __ballot(1);
        const int lane = threadIdx.x & (ops_detail::kWave - 1);
        const int leader = __ffsll((unsigned long long)active) - 1;
        unsigned long long base = 0;
        if (lane == leader) base = atomicAdd(end, (unsigned long long)__popcll(active));
        base = __shfl(base, leader, ops_detail::kWave);
        const uint64_t below = lane ? (active & (~0ull >> (64 - lane))) : 0ull;
        items[base + __popcll(below)] = v;
    }
};

// Host side of the queue: a device array of capacity maxLen and a device end
// counter; [curr, end) is the active part (getQueueAtCurr /
// getActiveQueueSize).  A full queue is the caller's bound to keep: with at
// most maxLen distinct enqueues (a visited check per vertex) it never fills.
class vertexQueue {
public:
    vertexQueue() = default;
    vertexQueue(const vertexQueue &) = delete;
    vertexQueue &operator=(const vertexQueue &) = delete;
    ~vertexQueue() { Free(); }

    int Init(int64_t maxLen)
    {
        Free();
        if (hipMalloc(&items_, (size_t)(maxLen > 0 ? maxLen : 1) * sizeof(int32_t)) != hipSuccess ||
            hipMalloc(&end_, sizeof(unsigned long long)) != hipSuccess) {
            Free();
            return TC_ERR_OOM;
        }
        maxLen_ = maxLen;
        return resetQueue();
    }
    void Free()
    {
        if (items_) (void)hipFree(items_);
        if (end_) (void)hipFree(end_);
        items_ = nullptr;
        end_ = nullptr;
        curr_ = 0;
    }
    int resetQueue(hipStream_t s = nullptr)
    {
        curr_ = 0;
        const unsigned long long z = 0;
        if (hipMemcpyAsync(end_, &z, sizeof(z), hipMemcpyHostToDevice, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            return TC_ERR_HIP;
        return TC_OK;
    }
    int enqueueFromHost(int32_t v, hipStream_t s = nullptr)
    {
        const int64_t e = getQueueEnd(s);
        if (e < 0 || e >= maxLen_) return TC_ERR_INVALID;
        const unsigned long long ne = (unsigned long long)e + 1;
        if (hipMemcpyAsync(items_ + e, &v, sizeof(v), hipMemcpyHostToDevice, s) != hipSuccess ||
            hipMemcpyAsync(end_, &ne, sizeof(ne), hipMemcpyHostToDevice, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            return TC_ERR_HIP;
        return TC_OK;
    }
    // The end counter as the device holds it after the stream's work (-1 on error).
    int64_t getQueueEnd(hipStream_t s = nullptr) const
    {
        unsigned long long e = 0;
        if (hipMemcpyAsync(&e, end_, sizeof(e), hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            return -1;
        return (int64_t)e;
    }
    int64_t getQueueCurr() const { return curr_; }
    void setQueueCurr(int64_t c) { curr_ = c; }
    int64_t getActiveQueueSize(hipStream_t s = nullptr) const { return getQueueEnd(s) - curr_; }
    int32_t *getQueue() const { return items_; }
    int32_t *getQueueAtCurr() const { return items_ + curr_; }
    DeviceVertexQueue device() const { return DeviceVertexQueue{items_, end_}; }

private:
    int32_t *items_ = nullptr;
    unsigned long long *end_ = nullptr;
    int64_t curr_ = 0;
    int64_t maxLen_ = 0;
};

// ------------------------------------------------------------- functors --

// op(view, v) for every vertex (verts null) or every entry of verts.
template <class Op> int for_vertices(tc_graph *g, const int32_t *verts, int64_t n, Op op)
{
    GraphView v;
    hipStream_t s;
    const int st = ops_detail::view_of(g, v, s);
    if (st != TC_OK) return st;
    if (!verts) n = v.nv;
    if (n <= 0) return TC_OK;
    hipLaunchKernelGGL(ops_detail::k_vertices<Op>, dim3(ops_detail::grid(n, ops_detail::kThreads)),
                       dim3(ops_detail::kThreads), 0, s, v, verts, n, op);
    return hipGetLastError() == hipSuccess ? TC_OK : TC_ERR_HIP;
}

// op(view, src, dst) for every edge of the vertices lb was prepared for.
template <class Op> int for_edges_lb(tc_graph *g, const LoadBalance &lb, Op op)
{
    GraphView v;
    hipStream_t s;
    const int st = ops_detail::view_of(g, v, s);
    if (st != TC_OK) return st;
    if (lb.size() <= 0) return TC_OK;
    // the edge count stays on the device: size the grid by the vertices' share
    // of the store and let the kernel stride over whatever the prefix holds
    const int64_t est = v.nnz > 0 ? v.nnz : 1;
    hipLaunchKernelGGL(ops_detail::k_edges_lb<Op>,
                       dim3(ops_detail::grid(est, ops_detail::kThreads)), dim3(ops_detail::kThreads),
                       0, s, v, lb.vertices(), lb.size(), lb.prefix(), op);
    return hipGetLastError() == hipSuccess ? TC_OK : TC_ERR_HIP;
}

// op(view, src[i], dst[i]) for every pair of a batch (device arrays).
template <class Op>
int for_batch_edges(tc_graph *g, const int32_t *src, const int32_t *dst, int64_t n, Op op)
{
    GraphView v;
    hipStream_t s;
    const int st = ops_detail::view_of(g, v, s);
    if (st != TC_OK) return st;
    if (n <= 0) return TC_OK;
    hipLaunchKernelGGL(ops_detail::k_batch_edges<Op>, dim3(ops_detail::grid(n, ops_detail::kThreads)),
                       dim3(ops_detail::kThreads), 0, s, v, src, dst, n, op);
    return hipGetLastError() == hipSuccess ? TC_OK : TC_ERR_HIP;
}

// ----------------------------------------- the reference's operator names --

// allVinG_TraverseVertices (operators.cuh:36-60).
template <cusSubKernelVertex F> int allVinG_TraverseVertices(tc_graph *g, void *metadata)
{
    return for_vertices(g, nullptr, 0, ops_detail::VertexFn<F>{metadata});
}

// allVinA_TraverseVertices (operators.cuh:93-125).
template <cusSubKernelVertex F>
int allVinA_TraverseVertices(tc_graph *g, void *metadata, const int32_t *verArray, int64_t len)
{
    if (!verArray) return len == 0 ? TC_OK : TC_ERR_INVALID;
    return for_vertices(g, verArray, len, ops_detail::VertexFn<F>{metadata});
}

// allVinA_TraverseEdges_LB (operators.cuh:291-310), with lb prepared for the
// array (the reference's needLoadBalance = false form when it already is).
template <cusSubKernelEdge F>
int allVinA_TraverseEdges_LB(tc_graph *g, void *metadata, LoadBalance &lb, const int32_t *verArray,
                             int64_t len, bool needLoadBalance = true)
{
    if (needLoadBalance) {
        const int st = lb.Prepare(g, verArray, len);
        if (st != TC_OK) return st;
    }
    return for_edges_lb(g, lb, ops_detail::EdgeFn<F>{metadata});
}

// allVinA_TraverseEdges (operators.cuh:131-160): the same traversal, with
// scratch of its own.
template <cusSubKernelEdge F>
int allVinA_TraverseEdges(tc_graph *g, void *metadata, const int32_t *verArray, int64_t len)
{
    if (!verArray) return len == 0 ? TC_OK : TC_ERR_INVALID;
    LoadBalance lb;
    const int st = allVinA_TraverseEdges_LB<F>(g, metadata, lb, verArray, len);
    (void)hipDeviceSynchronize();  // lb's scratch is freed on return
    return st;
}

// allEinG_TraverseEdges (operators.cuh:62-91): every edge of the store.
template <cusSubKernelEdge F> int allEinG_TraverseEdges(tc_graph *g, void *metadata)
{
    LoadBalance lb;
    const int st = allVinA_TraverseEdges_LB<F>(g, metadata, lb, nullptr, 0);
    (void)hipDeviceSynchronize();
    return st;
}

// allEinA_TraverseEdges (operators.cuh:315-344): every pair of a batch.
template <cusSubKernelEdge F>
int allEinA_TraverseEdges(tc_graph *g, void *metadata, const int32_t *src, const int32_t *dst,
                          int64_t n)
{
    return for_batch_edges(g, src, dst, n, ops_detail::EdgeFn<F>{metadata});
}

// The reference's batch form (operators.cuh:333-344): the size from
// bu.getHostBUD()->getBatchSize(), the device entries from
// bu.getDeviceBUD()->getSrc()/getDst() (a BatchUpdate of custinger_tc.hpp).
template <cusSubKernelEdge F, class Batch>
int allEinA_TraverseEdges(tc_graph *g, void *metadata, Batch &bu)
{
    const int64_t n = *(bu.getHostBUD()->getBatchSize());
    if (n == 0) return TC_OK;
    const auto *d = bu.getDeviceBUD();
    return for_batch_edges(g, d->getSrc(), d->getDst(), n, ops_detail::EdgeFn<F>{metadata});
}

}  // namespace cuStingerAlgs
