// ops_test.hip -- test library for include/custinger_ops.hpp: each entry
// point runs one operator with a small __device__ function or functor and
// leaves its result in caller-supplied device memory (checked from Python
// against numpy on the same CSR).
#include <climits>

#include "custinger_ops.hpp"

using namespace cuStingerAlgs;

namespace {

__device__ void op_degree(const GraphView &g, int32_t v, void *m)
{
    static_cast<uint32_t *>(m)[v] = (uint32_t)g.degree(v);
}

struct EdgeSums {
    unsigned long long *nbr_sum;
    uint32_t *cnt;
};
__device__ void op_edge_sums(const GraphView &, int32_t s, int32_t d, void *m)
{
    const EdgeSums *e = static_cast<const EdgeSums *>(m);
    atomicAdd(&e->nbr_sum[s], (unsigned long long)d);
    atomicAdd(&e->cnt[s], 1u);
}

__device__ void op_indegree(const GraphView &, int32_t, int32_t d, void *m)
{
    atomicAdd(&static_cast<uint32_t *>(m)[d], 1u);
}

__device__ void op_has_edge(const GraphView &g, int32_t s, int32_t d, void *m)
{
    const int32_t *a = g.adj(s);
    int32_t lo = 0, hi = g.degree(s);
    while (lo < hi) {
        const int32_t mid = (lo + hi) >> 1;
        if (a[mid] < d) lo = mid + 1; else hi = mid;
    }
    if (lo < g.degree(s) && a[lo] == d) atomicAdd(static_cast<unsigned long long *>(m), 1ull);
}

struct Bfs {
    int32_t *level;
    int32_t next;
    DeviceVertexQueue q;
};
__device__ void op_bfs_expand(const GraphView &, int32_t, int32_t d, void *m)
{
    const Bfs *b = static_cast<const Bfs *>(m);
    if (atomicCAS(&b->level[d], INT_MAX, b->next) == INT_MAX) b->q.enqueue(d);
}

// functor form: level[v] = INT_MAX
struct SetInfinity {
    int32_t *level;
    __device__ void operator()(const GraphView &, int32_t v) const { level[v] = INT_MAX; }
};

template <class T> T *to_device(const T &x)
{
    T *p = nullptr;
    if (hipMalloc(&p, sizeof(T)) != hipSuccess) return nullptr;
    if (hipMemcpy(p, &x, sizeof(T), hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(p);
        return nullptr;
    }
    return p;
}

int finish(int st)
{
    if (hipDeviceSynchronize() != hipSuccess) return TC_ERR_HIP;
    return st;
}

}  // namespace

extern "C" {

int ops_degrees(tc_graph *g, uint32_t *out_dev)
{
    return finish(allVinG_TraverseVertices<op_degree>(g, out_dev));
}

int ops_subset_degrees(tc_graph *g, const int32_t *verts_dev, int64_t n, uint32_t *out_dev)
{
    return finish(allVinA_TraverseVertices<op_degree>(g, out_dev, verts_dev, n));
}

int ops_edge_sums(tc_graph *g, const int32_t *verts_dev, int64_t n, unsigned long long *nbr_sum_dev,
                  uint32_t *cnt_dev)
{
    EdgeSums *m = to_device(EdgeSums{nbr_sum_dev, cnt_dev});
    if (!m) return TC_ERR_OOM;
    LoadBalance lb;
    int st = allVinA_TraverseEdges_LB<op_edge_sums>(g, m, lb, verts_dev, n);
    st = finish(st);
    (void)hipFree(m);
    return st;
}

int ops_edge_sums_unbalanced_name(tc_graph *g, const int32_t *verts_dev, int64_t n,
                                  unsigned long long *nbr_sum_dev, uint32_t *cnt_dev)
{
    EdgeSums *m = to_device(EdgeSums{nbr_sum_dev, cnt_dev});
    if (!m) return TC_ERR_OOM;
    int st = finish(allVinA_TraverseEdges<op_edge_sums>(g, m, verts_dev, n));
    (void)hipFree(m);
    return st;
}

int ops_indegree(tc_graph *g, uint32_t *out_dev)
{
    return finish(allEinG_TraverseEdges<op_indegree>(g, out_dev));
}

int ops_batch_found(tc_graph *g, const int32_t *src_dev, const int32_t *dst_dev, int64_t n,
                    unsigned long long *found_dev)
{
    return finish(allEinA_TraverseEdges<op_has_edge>(g, found_dev, src_dev, dst_dev, n));
}

// Top-down BFS from root (the reference's bfsTD shape,
// include/static_breadth_first_search/bfs_top_down.cuh:60-80): level[v] =
// hops from root, INT_MAX when unreached; *levels = number of frontiers.
int ops_bfs(tc_graph *g, int32_t root, int32_t *level_dev, int32_t *levels)
{
    tc_graph_view view;
    int st = tc_graph_get_view(g, &view);
    if (st != TC_OK) return st;
    if (root < 0 || root >= view.nv) return TC_ERR_INVALID;
    hipStream_t s = (hipStream_t)view.stream;
    st = for_vertices(g, nullptr, 0, SetInfinity{level_dev});
    if (st != TC_OK) return finish(st);
    const int32_t zero = 0;
    if (hipMemcpyAsync(level_dev + root, &zero, sizeof(zero), hipMemcpyHostToDevice, s) != hipSuccess)
        return finish(TC_ERR_HIP);
    vertexQueue q;
    if ((st = q.Init(view.nv)) != TC_OK || (st = q.enqueueFromHost(root, s)) != TC_OK)
        return finish(st);
    Bfs host{level_dev, 0, q.device()};
    Bfs *m = to_device(host);
    if (!m) return finish(TC_ERR_OOM);
    LoadBalance lb;
    int32_t depth = 0;
    for (;;) {
        const int64_t end = q.getQueueEnd(s);
        const int64_t active = end - q.getQueueCurr();
        if (end < 0) {
            st = TC_ERR_HIP;
            break;
        }
        if (active <= 0) break;
        ++depth;
        host.next = depth;
        if (hipMemcpyAsync(m, &host, sizeof(host), hipMemcpyHostToDevice, s) != hipSuccess) {
            st = TC_ERR_HIP;
            break;
        }
        st = allVinA_TraverseEdges_LB<op_bfs_expand>(g, m, lb, q.getQueueAtCurr(), active);
        if (st != TC_OK) break;
        q.setQueueCurr(end);
    }
    st = finish(st);
    (void)hipFree(m);
    *levels = depth;
    return st;
}

}  // extern "C"
