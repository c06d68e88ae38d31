// tc_capi.hip -- extern "C" entry points (include/custinger_tc.h).  No C++
// exception or HIP error crosses this boundary: every function returns a
// status and records a message for tc_last_error().
#include "tc_internal.hpp"

#include <cstring>
#include <exception>
#include <new>

struct tc_graph {
    tc::Graph *g;
};

namespace tc {
extern thread_local Error g_last;
int batch_apply(Graph &g, const int32_t *u, const int32_t *v, const uint8_t *is_del, int64_t B,
                uint32_t *delta, int64_t *dglobal, const tc_schedule &sched, tc_batch_stats *st,
                bool count, bool directed);
int graph_compare(Graph &a, Graph &b, int64_t *size_mismatch, int64_t *entry_mismatch,
                  int32_t *first_vertex);
int stream_delta_passes(Graph &g, int64_t nb, const int32_t *s_src, const int32_t *s_dst,
                        const int32_t *dup, const int64_t *boff, uint32_t *out, int32_t mask);
int batch_prepare(Graph &g, int64_t nb, const int32_t *src, const int32_t *dst, int32_t *s_src,
                  int32_t *s_dst, int32_t *dup, int64_t *bo);
int batched_count(int32_t G, const int64_t *voff, const int64_t *off, const int32_t *adj,
                  int64_t nnz, int32_t nv_total, uint64_t *per_graph, uint32_t *per_vertex,
                  const tc_schedule &sched, tc_stats *stats, hipStream_t s);
int lb_partition(Graph &g, const int32_t *verts, int64_t len, int64_t *parts, int64_t *P);
int lb_degree_bins(Graph &g, const int32_t *verts, int64_t len, int64_t counts[3]);
int lb_edge_prefix(Graph &g, const int32_t *verts, int64_t len, int64_t *pre);
int gen_rmat(int32_t scale, int64_t m, double a, double b, double c, uint64_t seed, int32_t *u,
             int32_t *v, hipStream_t s);
int gen_er(int32_t n, int64_t m, uint64_t seed, int32_t *u, int32_t *v, hipStream_t s);
int gen_rmat_multi(int32_t G, const int32_t *n, const int32_t *ef, const uint64_t *seeds, double a,
                   double b, double c, int32_t *u, int32_t *v, hipStream_t s);
int canon_csr_device(int32_t nv, int64_t m, const int32_t *u, const int32_t *v, int64_t **off,
                     int32_t **adj, int64_t *nnz, hipStream_t s);
int collect_timings(Graph &g, int64_t max_samples, uint64_t seed, float *feat, float *cost,
                    int64_t *n);
int mlp_train_step(float *params, float *m, float *v, int32_t step, const float *feat,
                   const float *target, int64_t n, float lr, float *grads, float *loss);
int mlp_forward(const float *params, const float *feat, int64_t n, float *pred);
int model_load(const float *params_host);
int model_table(uint8_t *table, int32_t *dim);
int strategy_table_set(const uint8_t *table, int32_t dim);
int autotune(Graph &g, tc_schedule *best, double *best_ms);
int batched_graph_create(int32_t G, const int64_t *voff, const int64_t *off, const int32_t *adj,
                         int64_t nnz, int32_t nv_total, hipStream_t s, Graph **out);
int batched_graph_count(Graph &g, uint64_t *per_graph, uint32_t *per_vertex,
                        const tc_schedule &sched, tc_stats *stats);
int set_default_schedule(const tc_schedule *s);
bool schedule_override(tc_schedule *s);
}  // namespace tc

#define TC_GUARD(...)                                                                  \
    try {                                                                              \
        __VA_ARGS__                                                                    \
    } catch (const std::bad_alloc &) {                                                 \
        tc::set_last_error(TC_ERR_OOM, "host allocation failed");                      \
        return TC_ERR_OOM;                                                             \
    } catch (const std::exception &ex) {                                               \
        tc::set_last_error(TC_ERR_INVALID, ex.what());                                 \
        return TC_ERR_INVALID;                                                         \
    } catch (...) {                                                                    \
        tc::set_last_error(TC_ERR_INVALID, "unknown exception");                       \
        return TC_ERR_INVALID;                                                         \
    }

#define TC_CHECK_ARG(cond, msg)                                                        \
    do {                                                                               \
        if (!(cond)) {                                                                 \
            tc::set_last_error(TC_ERR_INVALID, msg);                                   \
            return TC_ERR_INVALID;                                                     \
        }                                                                              \
    } while (0)

extern "C" {

const char *tc_last_error(void) { return tc::g_last.msg.c_str(); }

const char *tc_status_string(int status)
{
    switch (status) {
    case TC_OK: return "ok";
    case TC_ERR_INVALID: return "invalid argument";
    case TC_ERR_OOM: return "out of device memory";
    case TC_ERR_HIP: return "HIP runtime error";
    case TC_ERR_UNSUPPORTED: return "unsupported";
    case TC_ERR_CAPACITY: return "capacity exceeded";
    default: return "unknown status";
    }
}

int tc_version(void) { return 10000; }

int tc_device_alloc(void **ptr, uint64_t bytes)
{
    TC_CHECK_ARG(ptr, "tc_device_alloc: null ptr");
    if (hipMalloc(ptr, bytes ? bytes : 16) != hipSuccess) {
        (void)hipGetLastError();
        *ptr = nullptr;
        tc::set_last_error(TC_ERR_OOM, "hipMalloc failed");
        return TC_ERR_OOM;
    }
    return TC_OK;
}

int tc_device_free(void *ptr)
{
    if (!ptr) return TC_OK;
    TC_HIP(hipFree(ptr));
    return TC_OK;
}

int tc_memcpy(void *dst, const void *src, uint64_t bytes, int kind)
{
    if (!bytes) return TC_OK;
    TC_CHECK_ARG(dst && src, "tc_memcpy: null pointer");
    hipMemcpyKind k;
    switch (kind) {
    case 0: k = hipMemcpyHostToDevice; break;
    case 1: k = hipMemcpyDeviceToHost; break;
    case 2: k = hipMemcpyDeviceToDevice; break;
    case 3: std::memcpy(dst, src, bytes); return TC_OK;
    default: TC_CHECK_ARG(false, "tc_memcpy: bad kind");
    }
    TC_HIP(hipMemcpy(dst, src, bytes, k));
    return TC_OK;
}

int tc_memset(void *dst, int value, uint64_t bytes)
{
    if (!bytes) return TC_OK;
    TC_CHECK_ARG(dst, "tc_memset: null pointer");
    TC_HIP(hipMemset(dst, value, bytes));
    return TC_OK;
}

int tc_device_synchronize(void)
{
    TC_HIP(hipDeviceSynchronize());
    return TC_OK;
}

int tc_device_name(char *buf, int len)
{
    TC_CHECK_ARG(buf && len > 0, "tc_device_name: bad buffer");
    int dev = 0;
    TC_HIP(hipGetDevice(&dev));
    hipDeviceProp_t p;
    TC_HIP(hipGetDeviceProperties(&p, dev));
    std::snprintf(buf, (size_t)len, "%s (%s, %d CUs)", p.name, p.gcnArchName, p.multiProcessorCount);
    return TC_OK;
}

int tc_graph_create_csr(int32_t nv, int64_t nnz, const int64_t *off, const int32_t *adj,
                        int location, uint32_t flags, int32_t slack_pct, tc_graph **out)
{
    TC_GUARD({
        TC_CHECK_ARG(out, "tc_graph_create_csr: null out");
        *out = nullptr;
        tc::Graph *g = nullptr;
        int st = tc::graph_create_csr(nv, nnz, off, adj, location, flags, slack_pct, &g);
        if (st != TC_OK) return st;
        *out = new tc_graph{g};
        return TC_OK;
    })
}

int tc_graph_destroy(tc_graph *g)
{
    TC_GUARD({
        if (!g) return TC_OK;
        tc::graph_destroy(g->g);
        delete g;
        return TC_OK;
    })
}

int tc_graph_set_stream(tc_graph *g, void *hip_stream)
{
    TC_CHECK_ARG(g, "tc_graph_set_stream: null graph");
    g->g->stream = (hipStream_t)hip_stream;
    return TC_OK;
}

int tc_graph_get_info(tc_graph *g, tc_graph_info *info)
{
    TC_GUARD({
        TC_CHECK_ARG(g && info, "tc_graph_get_info: null argument");
        tc::Graph &G = *g->g;
        if (G.symmetric < 0 || G.self_loops < 0) {
            int64_t bad = 0;
            TC_TRY(tc::validate_graph(G, true, &bad));
        }
        std::memset(info, 0, sizeof(*info));
        info->nv = G.st.nv;
        info->nnz = G.st.nnz;
        info->allocated = G.st.pool_used;
        info->symmetric = G.symmetric;
        info->self_loops = G.self_loops;
        info->max_degree = G.max_degree;
        info->reallocations = G.st.reallocs;
        return TC_OK;
    })
}

int tc_graph_to_csr(tc_graph *g, int64_t *off, int32_t *adj, int location)
{
    TC_GUARD({
        TC_CHECK_ARG(g && off && (adj || g->g->st.nnz == 0), "tc_graph_to_csr: null argument");
        TC_CHECK_ARG(location == TC_HOST || location == TC_DEVICE, "tc_graph_to_csr: bad location");
        return tc::graph_to_csr(*g->g, off, adj, location);
    })
}

int tc_graph_degrees(tc_graph *g, int32_t *deg_dev)
{
    TC_CHECK_ARG(g && deg_dev, "tc_graph_degrees: null argument");
    if (g->g->st.nv == 0) return TC_OK;
    TC_HIP(hipMemcpyAsync(deg_dev, g->g->st.len.p, (size_t)g->g->st.nv * sizeof(int32_t),
                          hipMemcpyDeviceToDevice, g->g->stream));
    TC_HIP(hipStreamSynchronize(g->g->stream));
    return TC_OK;
}

int tc_schedule_default(tc_schedule *s)
{
    TC_CHECK_ARG(s, "tc_schedule_default: null");
    *s = tc::default_schedule();
    tc::schedule_override(s);
    return TC_OK;
}

static int check_schedule(const tc_schedule *s)
{
    if (!s) return TC_OK;
    TC_CHECK_ARG(s->method >= TC_METHOD_AUTO && s->method <= TC_METHOD_DIRECT,
                 "tc_schedule: bad method");
    TC_CHECK_ARG(s->policy >= TC_POLICY_FIXED && s->policy <= TC_POLICY_MERGEPATH,
                 "tc_schedule: bad policy");
    TC_CHECK_ARG(s->small_max >= 1 && s->mid_max >= s->small_max && s->job_items >= 1,
                 "tc_schedule: bad bin thresholds");
    TC_CHECK_ARG(s->build == 0 || s->build == 1, "tc_schedule: bad build");
    return TC_OK;
}

int tc_static_count(tc_graph *g, uint32_t *per_vertex_dev, uint64_t *global_host,
                    const tc_schedule *sched, tc_stats *stats)
{
    TC_GUARD({
        TC_CHECK_ARG(g, "tc_static_count: null graph");
        TC_TRY(check_schedule(sched));
        tc_schedule s;
        if (sched) s = *sched; else tc_schedule_default(&s);
        return tc::static_count(*g->g, per_vertex_dev, global_host, s, stats, false);
    })
}

int tc_csr_count_device(int32_t nv, int64_t nnz, const int64_t *d_off, const int32_t *d_ind,
                        uint32_t *out_dev, uint64_t *global_host, void *hip_stream)
{
    TC_GUARD({
        TC_CHECK_ARG(nv >= 0 && nnz >= 0 && (nv == 0 || d_off) && (nnz == 0 || d_ind),
                     "tc_csr_count_device: bad arguments");
        tc::Graph *g = nullptr;
        TC_TRY(tc::graph_create_csr(nv, nnz, d_off, d_ind, TC_DEVICE, TC_GRAPH_ASSUME_SYMMETRIC, 0,
                                    &g));
        g->stream = (hipStream_t)hip_stream;
        tc_schedule s = tc::default_schedule();
        s.method = TC_METHOD_DIRECT;
        int st = tc::static_count(*g, out_dev, global_host, s, nullptr, true);
        tc::graph_destroy(g);
        return st;
    })
}

int tc_csr_count_device32(int32_t nv, const int32_t *d_off, const int32_t *d_ind, uint32_t *out_dev,
                          uint64_t *global_host, void *hip_stream)
{
    TC_GUARD({
        TC_CHECK_ARG(nv >= 0 && (nv == 0 || d_off), "tc_csr_count_device32: bad arguments");
        int32_t nnz = 0;
        if (nv > 0) {
            TC_HIP(hipMemcpyAsync(&nnz, d_off + nv, sizeof(nnz), hipMemcpyDeviceToHost,
                                  (hipStream_t)hip_stream));
            TC_HIP(hipStreamSynchronize((hipStream_t)hip_stream));
        }
        TC_CHECK_ARG(nnz >= 0 && (nnz == 0 || d_ind), "tc_csr_count_device32: bad arguments");
        tc::Graph *g = nullptr;
        TC_TRY(tc::graph_create_csr(nv, nnz, reinterpret_cast<const int64_t *>(d_off), d_ind,
                                    TC_DEVICE, TC_GRAPH_ASSUME_SYMMETRIC | TC_GRAPH_OFFSETS_I32, 0, &g));
        g->stream = (hipStream_t)hip_stream;
        tc_schedule s = tc::default_schedule();
        s.method = TC_METHOD_DIRECT;
        int st = tc::static_count(*g, out_dev, global_host, s, nullptr, true);
        tc::graph_destroy(g);
        return st;
    })
}

int tc_batch_apply(tc_graph *g, const int32_t *u, const int32_t *v, const uint8_t *is_delete,
                   int64_t B, uint32_t *per_vertex_delta_dev, int64_t *delta_global_host,
                   const tc_schedule *sched, tc_batch_stats *stats)
{
    TC_GUARD({
        TC_CHECK_ARG(g && B >= 0 && (B == 0 || (u && v)), "tc_batch_apply: bad arguments");
        TC_TRY(check_schedule(sched));
        tc_schedule s;
        if (sched) s = *sched; else tc_schedule_default(&s);
        return tc::batch_apply(*g->g, u, v, is_delete, B, per_vertex_delta_dev, delta_global_host,
                               s, stats, true, false);
    })
}

int tc_batch_update_only(tc_graph *g, const int32_t *u, const int32_t *v,
                         const uint8_t *is_delete, int64_t B, tc_batch_stats *stats)
{
    TC_GUARD({
        TC_CHECK_ARG(g && B >= 0 && (B == 0 || (u && v)), "tc_batch_update_only: bad arguments");
        tc_schedule s = tc::default_schedule();
        return tc::batch_apply(*g->g, u, v, is_delete, B, nullptr, nullptr, s, stats, false, false);
    })
}

int tc_batch_update_directed(tc_graph *g, const int32_t *src, const int32_t *dst,
                             const uint8_t *is_delete, int64_t B, tc_batch_stats *stats)
{
    TC_GUARD({
        TC_CHECK_ARG(g && B >= 0 && (B == 0 || (src && dst)),
                     "tc_batch_update_directed: bad arguments");
        tc_schedule s = tc::default_schedule();
        return tc::batch_apply(*g->g, src, dst, is_delete, B, nullptr, nullptr, s, stats, false, true);
    })
}

int tc_graph_set_growth(tc_graph *g, int32_t mul_x16, int32_t add)
{
    TC_GUARD({
        TC_CHECK_ARG(g && mul_x16 >= 0 && add >= 0 && (mul_x16 == 0 || mul_x16 >= 16 || add > 0),
                     "tc_graph_set_growth: bad policy (mul_x16 >= 16, or a positive add)");
        g->g->grow_mul_x16 = mul_x16;
        g->g->grow_add = add;
        return TC_OK;
    })
}

int tc_graph_compare(tc_graph *a, tc_graph *b, int64_t *size_mismatch, int64_t *entry_mismatch,
                     int32_t *first_vertex)
{
    TC_GUARD({
        TC_CHECK_ARG(a && b && size_mismatch && entry_mismatch && first_vertex,
                     "tc_graph_compare: bad arguments");
        return tc::graph_compare(*a->g, *b->g, size_mismatch, entry_mismatch, first_vertex);
    })
}

int tc_batch_prepare(tc_graph *g, int64_t nb, const int32_t *src, const int32_t *dst,
                     int32_t *s_src, int32_t *s_dst, int32_t *dup, int64_t *boff)
{
    TC_GUARD({
        TC_CHECK_ARG(g && nb >= 0 && (nb == 0 || (src && dst && s_src && s_dst && dup)) && boff,
                     "tc_batch_prepare: bad arguments");
        return tc::batch_prepare(*g->g, nb, src, dst, s_src, s_dst, dup, boff);
    })
}

int tc_stream_delta_passes(tc_graph *g, int64_t nb, const int32_t *s_src, const int32_t *s_dst,
                           const int32_t *dup, const int64_t *boff, uint32_t *out_dev,
                           int32_t pass_mask)
{
    TC_GUARD({
        TC_CHECK_ARG(g && nb >= 0 && (nb == 0 || (s_src && s_dst && dup)) && boff && out_dev,
                     "tc_stream_delta_passes: bad arguments");
        return tc::stream_delta_passes(*g->g, nb, s_src, s_dst, dup, boff, out_dev, pass_mask);
    })
}

int tc_batched_count(int32_t G, const int64_t *voff_dev, const int64_t *off_dev,
                     const int32_t *adj_dev, int64_t nnz, int32_t nv_total,
                     uint64_t *per_graph_dev, uint32_t *per_vertex_dev, const tc_schedule *sched,
                     tc_stats *stats, void *hip_stream)
{
    TC_GUARD({
        TC_CHECK_ARG(G >= 0 && voff_dev && off_dev && (nnz == 0 || adj_dev) && per_graph_dev,
                     "tc_batched_count: bad arguments");
        TC_TRY(check_schedule(sched));
        tc_schedule s;
        if (sched) s = *sched; else tc_schedule_default(&s);
        return tc::batched_count(G, voff_dev, off_dev, adj_dev, nnz, nv_total, per_graph_dev,
                                 per_vertex_dev, s, stats, (hipStream_t)hip_stream);
    })
}

int tc_batched_graph_create(int32_t G, const int64_t *voff_dev, const int64_t *off_dev,
                            const int32_t *adj_dev, int64_t nnz, int32_t nv_total, void *hip_stream,
                            tc_graph **out)
{
    TC_GUARD({
        TC_CHECK_ARG(out && G > 0 && nv_total >= 0 && nnz >= 0 && voff_dev && off_dev &&
                         (nnz == 0 || adj_dev),
                     "tc_batched_graph_create: bad arguments");
        *out = nullptr;
        tc::Graph *g = nullptr;
        TC_TRY(tc::batched_graph_create(G, voff_dev, off_dev, adj_dev, nnz, nv_total,
                                        (hipStream_t)hip_stream, &g));
        *out = new tc_graph{g};
        return TC_OK;
    })
}

int tc_batched_graph_count(tc_graph *g, uint64_t *per_graph_dev, uint32_t *per_vertex_dev,
                           const tc_schedule *sched, tc_stats *stats)
{
    TC_GUARD({
        TC_CHECK_ARG(g && per_graph_dev, "tc_batched_graph_count: bad arguments");
        TC_TRY(check_schedule(sched));
        tc_schedule s;
        if (sched) s = *sched; else tc_schedule_default(&s);
        return tc::batched_graph_count(*g->g, per_graph_dev, per_vertex_dev, s, stats);
    })
}

int tc_lb_partition(tc_graph *g, const int32_t *verts_dev, int64_t len, int64_t *parts_dev,
                    int64_t *P)
{
    TC_GUARD({
        TC_CHECK_ARG(g && verts_dev && len > 0 && parts_dev && P, "tc_lb_partition: bad arguments");
        return tc::lb_partition(*g->g, verts_dev, len, parts_dev, P);
    })
}

int tc_lb_edge_prefix(tc_graph *g, const int32_t *verts_dev, int64_t len, int64_t *pre_dev)
{
    TC_GUARD({
        TC_CHECK_ARG(g && pre_dev, "tc_lb_edge_prefix: null argument");
        if (!verts_dev) len = g->g->st.nv;
        TC_CHECK_ARG(len >= 0, "tc_lb_edge_prefix: bad length");
        return tc::lb_edge_prefix(*g->g, verts_dev, len, pre_dev);
    })
}

int tc_graph_get_view(tc_graph *g, tc_graph_view *view)
{
    TC_GUARD({
        TC_CHECK_ARG(g && view, "tc_graph_get_view: null argument");
        TC_CHECK_ARG(!g->g->batch_voff, "tc_graph_get_view: not available on a batched handle");
        const tc::Graph &G = *g->g;
        std::memset(view, 0, sizeof(*view));
        view->nv = G.st.nv;
        view->nnz = G.st.nnz;
        view->boff = G.st.boff.as<int64_t>();
        view->len = G.st.len.as<int32_t>();
        view->pool = G.st.pool.as<int32_t>();
        view->stream = (void *)G.stream;
        return TC_OK;
    })
}

int tc_lb_degree_bins(tc_graph *g, const int32_t *verts_dev, int64_t len, int64_t counts[3])
{
    TC_GUARD({
        TC_CHECK_ARG(g && (verts_dev || len == 0) && len >= 0 && counts,
                     "tc_lb_degree_bins: bad arguments");
        return tc::lb_degree_bins(*g->g, verts_dev, len, counts);
    })
}

int tc_gen_rmat_pairs(int32_t scale, int64_t m, double a, double b, double c, uint64_t seed,
                      int32_t *u_dev, int32_t *v_dev, void *hip_stream)
{
    TC_GUARD({
        TC_CHECK_ARG(scale >= 1 && scale <= 30 && m >= 0 && (m == 0 || (u_dev && v_dev)) && a >= 0 &&
                         b >= 0 && c >= 0 && a + b + c <= 1.0,
                     "tc_gen_rmat_pairs: bad arguments");
        return tc::gen_rmat(scale, m, a, b, c, seed, u_dev, v_dev, (hipStream_t)hip_stream);
    })
}

int tc_gen_er_pairs(int32_t n, int64_t m, uint64_t seed, int32_t *u_dev, int32_t *v_dev,
                    void *hip_stream)
{
    TC_GUARD({
        TC_CHECK_ARG(n >= 1 && m >= 0 && (m == 0 || (u_dev && v_dev)), "tc_gen_er_pairs: bad arguments");
        return tc::gen_er(n, m, seed, u_dev, v_dev, (hipStream_t)hip_stream);
    })
}

int tc_gen_rmat_multi(int32_t G, const int32_t *n, const int32_t *ef, const uint64_t *seeds,
                      double a, double b, double c, int32_t *u_dev, int32_t *v_dev, void *hip_stream)
{
    TC_GUARD({
        TC_CHECK_ARG(G >= 0 && (G == 0 || (n && ef && seeds && u_dev && v_dev)) && a >= 0 && b >= 0 &&
                         c >= 0 && a + b + c <= 1.0,
                     "tc_gen_rmat_multi: bad arguments");
        return tc::gen_rmat_multi(G, n, ef, seeds, a, b, c, u_dev, v_dev, (hipStream_t)hip_stream);
    })
}

int tc_canon_csr_device(int32_t nv, int64_t m, const int32_t *u_dev, const int32_t *v_dev,
                        int64_t **off_dev, int32_t **adj_dev, int64_t *nnz, void *hip_stream)
{
    TC_GUARD({
        TC_CHECK_ARG(nv >= 1 && m >= 0 && (m == 0 || (u_dev && v_dev)) && off_dev && adj_dev && nnz,
                     "tc_canon_csr_device: bad arguments");
        return tc::canon_csr_device(nv, m, u_dev, v_dev, off_dev, adj_dev, nnz,
                                    (hipStream_t)hip_stream);
    })
}

int tc_collect_timings(tc_graph *g, int64_t max_samples, uint64_t seed, float *feat_host,
                       float *cost_host, int64_t *n)
{
    TC_GUARD({
        TC_CHECK_ARG(g && max_samples > 0 && feat_host && cost_host && n,
                     "tc_collect_timings: bad arguments");
        return tc::collect_timings(*g->g, max_samples, seed, feat_host, cost_host, n);
    })
}

int tc_mlp_train_step(float *params_dev, float *adam_m_dev, float *adam_v_dev, int32_t step,
                      const float *feat_dev, const float *target_dev, int64_t n, float lr,
                      float *grads_dev, float *loss_host)
{
    TC_GUARD({
        TC_CHECK_ARG(params_dev && feat_dev && target_dev && n > 0 && step >= 1,
                     "tc_mlp_train_step: bad arguments");
        return tc::mlp_train_step(params_dev, adam_m_dev, adam_v_dev, step, feat_dev, target_dev,
                                  n, lr, grads_dev, loss_host);
    })
}

int tc_mlp_forward(const float *params_dev, const float *feat_dev, int64_t n, float *pred_dev)
{
    TC_GUARD({
        TC_CHECK_ARG(params_dev && feat_dev && pred_dev && n >= 0, "tc_mlp_forward: bad arguments");
        return tc::mlp_forward(params_dev, feat_dev, n, pred_dev);
    })
}

int tc_model_load(const float *params_host)
{
    TC_GUARD({
        return tc::model_load(params_host);
    })
}

int tc_strategy_table_set(const uint8_t *table_host, int32_t dim)
{
    TC_GUARD({
        return tc::strategy_table_set(table_host, dim);
    })
}

int tc_model_table(uint8_t *table_host, int32_t *dim)
{
    TC_GUARD({
        TC_CHECK_ARG(dim, "tc_model_table: null dim");
        return tc::model_table(table_host, dim);
    })
}

int tc_autotune(tc_graph *g, tc_schedule *best, double *best_ms)
{
    TC_GUARD({
        TC_CHECK_ARG(g && best, "tc_autotune: bad arguments");
        return tc::autotune(*g->g, best, best_ms);
    })
}

int tc_schedule_set_default(const tc_schedule *s)
{
    TC_GUARD({
        TC_TRY(check_schedule(s));
        return tc::set_default_schedule(s);
    })
}

}  // extern "C"
