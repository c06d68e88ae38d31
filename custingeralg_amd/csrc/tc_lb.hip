// tc_lb.hip -- the load-balance entry points and batched multi-graph counts.
//
// tc_lb_partition restates cusLoadBalance::LoadBalanceArray
// (include/load_balance.cuh:77-102, src/load_balance.cu:32-66) on the
// device: work = degree of each queued vertex, exclusive scan over len+1
// entries, P in {200, 500, 1000} by average work (capped at len), needles
// total*b/P computed in double precision and truncated, and lower-bound
// partition points over the first len-1 prefix entries.  The counting path
// itself uses the finer prefix-sum job partition of tc_static.hip.
#include "tc_internal.hpp"

#include <algorithm>

namespace tc {

namespace {

// work[i] = degree of verts[i] (of vertex i when verts is null), work[len] = 0
__global__ void k_lb_work(const int32_t *__restrict__ verts, int64_t len,
                          const int32_t *__restrict__ deg, int64_t *__restrict__ work)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i <= len;
         i += (int64_t)gridDim.x * blockDim.x)
        work[i] = i < len ? (int64_t)deg[verts ? verts[i] : (int32_t)i] : 0;
}

__global__ void k_lb_parts(const int64_t *__restrict__ prefix, int64_t len, int64_t P,
                           int64_t *__restrict__ parts)
{
    for (int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; b < P;
         b += (int64_t)gridDim.x * blockDim.x) {
        const int64_t needle = (int64_t)((double)prefix[len] * ((double)b / (double)P));
        int64_t lo = 0, hi = len - 1;
        while (lo < hi) {
            const int64_t mid = (lo + hi) >> 1;
            if (prefix[mid] < needle) lo = mid + 1; else hi = mid;
        }
        parts[b] = lo;
    }
}

__global__ void k_lb_bins(const int32_t *__restrict__ verts, int64_t len,
                          const int32_t *__restrict__ deg, unsigned long long *__restrict__ cnt)
{
    unsigned long long c[3] = {0, 0, 0};
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < len;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t d = deg[verts[i]];
        c[d >= 64 ? 0 : (d >= 16 ? 1 : 2)]++;
    }
    for (int q = 0; q < 3; ++q) {
        unsigned long long x = c[q];
        for (int dd = 32; dd > 0; dd >>= 1) x += __shfl_xor(x, dd);
        if ((threadIdx.x & 63) == 0 && x) atomicAdd(&cnt[q], x);
    }
}

__global__ void k_zero_ull(unsigned long long *p, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

}  // namespace

int lb_partition(Graph &g, const int32_t *verts, int64_t len, int64_t *parts, int64_t *P)
{
    hipStream_t s = g.stream;
    DevBuf work, prefix;
    auto done = [&](int st) {
        work.release();
        prefix.release();
        return st;
    };
    int st;
    if ((st = work.ensure((size_t)(len + 1) * sizeof(int64_t))) != TC_OK) return done(st);
    if ((st = prefix.ensure((size_t)(len + 1) * sizeof(int64_t))) != TC_OK) return done(st);
    hipLaunchKernelGGL(k_lb_work, dim3(grid_for(len + 1, 256, 65536)), dim3(256), 0, s, verts, len,
                       g.st.len.as<int32_t>(), work.as<int64_t>());
    if ((st = scan_exclusive_i64(work.as<int64_t>(), prefix.as<int64_t>(), len + 1, nullptr, g.tmp,
                                 s)) != TC_OK)
        return done(st);
    int64_t total = 0;
    if (hipMemcpyAsync(&total, prefix.as<int64_t>() + len, sizeof(total), hipMemcpyDeviceToHost, s) !=
            hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
        set_last_error(TC_ERR_HIP, "lb_partition: readback");
        return done(TC_ERR_HIP);
    }
    const float avg = (float)total / (float)len;
    int64_t np = avg < 4 ? 200 : (avg < 7 ? 500 : 1000);
    if (np > len) np = len;
    hipLaunchKernelGGL(k_lb_parts, dim3(grid_for(np, 256, 64)), dim3(256), 0, s, prefix.as<int64_t>(),
                       len, np, parts);
    if (hipStreamSynchronize(s) != hipSuccess || hipGetLastError() != hipSuccess) {
        set_last_error(TC_ERR_HIP, "lb_partition: kernels");
        return done(TC_ERR_HIP);
    }
    *P = np;
    return done(TC_OK);
}

int lb_edge_prefix(Graph &g, const int32_t *verts, int64_t len, int64_t *pre)
{
    hipLaunchKernelGGL(k_lb_work, dim3(grid_for(len + 1, 256, 65536)), dim3(256), 0, g.stream, verts,
                       len, g.st.len.as<int32_t>(), pre);
    TC_HIP(hipGetLastError());
    return scan_exclusive_i64(pre, pre, len + 1, nullptr, g.tmp, g.stream);
}

int lb_degree_bins(Graph &g, const int32_t *verts, int64_t len, int64_t counts[3])
{
    TC_TRY(g.accum.ensure(16 * sizeof(unsigned long long)));
    unsigned long long *acc = g.accum.as<unsigned long long>();
    hipLaunchKernelGGL(k_zero_ull, dim3(1), dim3(64), 0, g.stream, acc, (int64_t)3);
    if (len > 0)
        hipLaunchKernelGGL(k_lb_bins, dim3(grid_for(len, 256, 4096)), dim3(256), 0, g.stream, verts,
                           len, g.st.len.as<int32_t>(), acc);
    unsigned long long h[3];
    TC_HIP(hipMemcpyAsync(h, acc, sizeof(h), hipMemcpyDeviceToHost, g.stream));
    TC_HIP(hipStreamSynchronize(g.stream));
    for (int q = 0; q < 3; ++q) counts[q] = (int64_t)h[q];
    return TC_OK;
}

int batched_graph_create(int32_t G, const int64_t *voff, const int64_t *off, const int32_t *adj,
                         int64_t nnz, int32_t nv_total, hipStream_t s, Graph **out)
{
    *out = nullptr;
    Graph *g = nullptr;
    TC_TRY(graph_borrow_csr(nv_total, nnz, off, adj, s, &g));
    g->batch_voff = voff;
    g->batch_ngraphs = G;
    const int st = validate_graph(*g, true, nullptr);
    if (st != TC_OK) {
        graph_destroy(g);
        return st;
    }
    *out = g;
    return TC_OK;
}

int batched_graph_count(Graph &g, uint64_t *per_graph, uint32_t *per_vertex,
                        const tc_schedule &sched, tc_stats *stats)
{
    if (g.batch_ngraphs <= 0 || !g.batch_voff) {
        set_last_error(TC_ERR_INVALID, "tc_batched_graph_count: not a batched graph handle");
        return TC_ERR_INVALID;
    }
    hipLaunchKernelGGL(k_zero_ull, dim3(grid_for(g.batch_ngraphs, 256, 4096)), dim3(256), 0,
                       g.stream, reinterpret_cast<unsigned long long *>(per_graph),
                       (int64_t)g.batch_ngraphs);
    g.batch_per_graph = reinterpret_cast<unsigned long long *>(per_graph);
    uint64_t glob = 0;
    const int st = static_count(g, per_vertex, &glob, sched, stats, false);
    g.batch_per_graph = nullptr;
    return st;
}

int batched_count(int32_t G, const int64_t *voff, const int64_t *off, const int32_t *adj,
                  int64_t nnz, int32_t nv_total, uint64_t *per_graph, uint32_t *per_vertex,
                  const tc_schedule &sched, tc_stats *stats, hipStream_t s)
{
    if (G == 0) return TC_OK;
    Graph *g = nullptr;
    TC_TRY(batched_graph_create(G, voff, off, adj, nnz, nv_total, s, &g));
    const int st = batched_graph_count(*g, per_graph, per_vertex, sched, stats);
    graph_destroy(g);
    return st;
}

}  // namespace tc
