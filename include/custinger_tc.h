/*
 * custinger_tc.h -- C-ABI of the MI355X triangle-counting engine.
 *
 * Drop-in boundary for cuStingerAlg's static and streaming triangle counting.
 * Plain pointers and sizes only; every entry point returns an int status
 * (TC_OK = 0) and never throws.  Device pointers are HIP device addresses on
 * the current device; "host" arguments are ordinary host memory.
 *
 * Count convention (bit-exact with the reference):
 *   per-vertex  uint32 out[v] = sum over neighbours u != v with deg(u) >= 2 and
 *               deg(v) >= 2 of |N(v) cap N(u)|, modulo 2^32 -- for a simple
 *               undirected graph this is 2*t(v)
 *               (src/static_triangle_counting/cct.cu:233-268);
 *   global      uint64 = the exact (unwrapped) sum of the per-vertex values
 *               (6*T for a simple undirected graph; tests/cct.cpp:119-125 sums
 *               the same values into int64).
 *
 * Threading: a handle is used by one host thread at a time; all work of a
 * handle is ordered on its stream (tc_graph_set_stream).
 */
#ifndef CUSTINGER_TC_H
#define CUSTINGER_TC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------ status -- */
enum {
    TC_OK = 0,
    TC_ERR_INVALID = 1,   /* bad argument or malformed graph/batch        */
    TC_ERR_OOM = 2,       /* device allocation failed                     */
    TC_ERR_HIP = 3,       /* HIP runtime error (see tc_last_error)        */
    TC_ERR_UNSUPPORTED = 4,
    TC_ERR_CAPACITY = 5   /* input larger than the entry point's limit    */
};

/* Message of the last failing call on this thread (never NULL). */
const char *tc_last_error(void);
const char *tc_status_string(int status);
/* Library version as MAJOR*10000 + MINOR*100 + PATCH. */
int tc_version(void);

/* ------------------------------------------------ device memory utils -- */
/* allocDeviceArray / freeDeviceArray / copyArray{HostToDevice,DeviceToHost,
 * DeviceToDevice} of the cuStinger utilities (used at include/algs.cuh:41,73,101). */
int tc_device_alloc(void **ptr, uint64_t bytes);
int tc_device_free(void *ptr);
/* kind: 0 host->device, 1 device->host, 2 device->device, 3 host->host */
int tc_memcpy(void *dst, const void *src, uint64_t bytes, int kind);
int tc_memset(void *dst, int value, uint64_t bytes);
int tc_device_synchronize(void);
int tc_device_name(char *buf, int len);

/* ------------------------------------------------------- graph store -- */
typedef struct tc_graph tc_graph;

/* Location of CSR input arrays. */
#define TC_HOST 0
#define TC_DEVICE 1

/* Create flags. */
#define TC_GRAPH_ASSUME_SYMMETRIC 0x1u /* skip the symmetry check (caller vouches) */
#define TC_GRAPH_NO_VALIDATE      0x2u /* skip sortedness/range validation        */
#define TC_GRAPH_OFFSETS_I32      0x4u /* off is int32[nv+1] (the reference's length_t) */

/* cuStinger::initializeCuStinger with eInitStateCSR (tests/cct.cpp:261-277):
 * build the dynamic per-vertex block store from a CSR (off[nv+1] int64,
 * adj[nnz] int32; every row sorted strictly ascending, ids in [0,nv)).
 * slack_pct adds capacity headroom per block for later insertions
 * (0 = exact fit; blocks still round up to a power of two when they grow). */
int tc_graph_create_csr(int32_t nv, int64_t nnz, const int64_t *off, const int32_t *adj,
                        int location, uint32_t flags, int32_t slack_pct, tc_graph **out);
/* cuStinger::freecuStinger (tests/cct.cpp:281). */
int tc_graph_destroy(tc_graph *g);
/* Bind the handle to a HIP stream (NULL = the default stream). */
int tc_graph_set_stream(tc_graph *g, void *hip_stream);

typedef struct {
    int32_t nv;
    int64_t nnz;               /* entries used (getNumberEdgesUsed, tests/cct_upd.cpp:127) */
    int64_t allocated;         /* entries allocated (getNumberEdgesAllocated, :128)       */
    int32_t symmetric;         /* 1 if every (u,v) has its (v,u)                         */
    int64_t self_loops;        /* number of (v,v) entries                                 */
    int32_t max_degree;
    int64_t reallocations;     /* block moves since creation ("allocs", :392,398)         */
} tc_graph_info;
int tc_graph_get_info(tc_graph *g, tc_graph_info *info);

/* Block growth policy, the update allocator of cuStinger(initAlloc,
 * updateAlloc) (tests/cct.cpp:261 passes defaultUpdateAllocater, 2x): a row
 * that outgrows its block moves to a block of len * mul_x16 / 16 + add
 * entries (at least len).  mul_x16 = add = 0 restores the default, the next
 * power of two. */
int tc_graph_set_growth(tc_graph *g, int32_t mul_x16, int32_t add);

/* compareCUS (src/streaming_triangle_counting/cct_upd.cu:442-484): compares
 * two stores of the same vertex count row by row on the device.
 * size_mismatch receives the rows whose used lengths differ, entry_mismatch
 * the rows of equal length whose entries differ, first_vertex the first
 * differing vertex (-1 if the stores are equal). */
int tc_graph_compare(tc_graph *a, tc_graph *b, int64_t *size_mismatch, int64_t *entry_mismatch,
                     int32_t *first_vertex);

/* Snapshot of the store as a compact CSR (location TC_HOST or TC_DEVICE);
 * off has nv+1 entries, adj room for info.nnz entries. */
int tc_graph_to_csr(tc_graph *g, int64_t *off, int32_t *adj, int location);
/* Per-vertex degrees (dVD->getUsed(), cct.cu:235) into a device int32[nv]. */
int tc_graph_degrees(tc_graph *g, int32_t *deg_dev);

/* ---------------------------------------------------------- schedule -- */
/* Strategy policy for the per-edge intersection. */
#define TC_POLICY_FIXED   0   /* degree-ratio heuristic                    */
#define TC_POLICY_LEARNED 1   /* MLP cost model table (tc_model_*)         */
#define TC_POLICY_MERGE   2   /* force one strategy (tests, timing)        */
#define TC_POLICY_BSEARCH 3
#define TC_POLICY_HASH    4
/* coalesced merge path: one wavefront splits one pair by merge-path
 * diagonals (singleIntersection / workPerThreadIP / bSearchIP /
 * fixStartPointIP / intersectPath, src/static_triangle_counting/cct.cu:54-199) */
#define TC_POLICY_MERGEPATH 5

/* Counting method. */
#define TC_METHOD_AUTO     0  /* oriented if symmetric and loop-free, else direct */
#define TC_METHOD_ORIENTED 1  /* degree-oriented, each triangle found once         */
#define TC_METHOD_DIRECT   2  /* every directed entry, like cct.cu:238-266         */

typedef struct {
    int32_t method;        /* TC_METHOD_*                                           */
    int32_t policy;        /* TC_POLICY_*                                           */
    int32_t small_max;     /* rows with |N+| <= small_max use one wave per job      */
    int32_t mid_max;       /* rows with |N+| <= mid_max use the mid bin              */
    int32_t job_items;     /* target items per job (prefix-sum partition grain)      */
    int32_t ratio_x16;     /* fixed heuristic: merge when long*16 <= short*ratio_x16 */
                           /*   and long <= 96 (0, the default: never merge)         */
    int32_t persistent;    /* 1: persistent job-queue launch                          */
    int32_t flush_items;   /* min items per owner slot in one job (bounds the per-job */
                           /*   flush of third-vertex counters)                      */
    int32_t hub_degree;    /* oriented: owners of at least this degree build their   */
                           /*   task lists from their own rows (0: default, none;    */
                           /*   65536 on graphs too large for the partitioned        */
                           /*   transpose)                                            */
    int32_t mpath_min;     /* fixed heuristic: balanced pairs (long <= 2 * short)     */
                           /*   whose shorter list has at least this many entries    */
                           /*   run as a merge path (0, the default: never)          */
    int32_t capture;       /* 1: repeated counts of an unchanged graph with the same  */
                           /*   schedule and output run as one hipGraph launch: the  */
                           /*   second such count records the whole sequence, later  */
                           /*   ones replay it (no host read-backs in between; the   */
                           /*   stats then give ms_total only, graph_replay = 1)     */
    int32_t build;         /* oriented layout: 0 (default) the partitioned transpose  */
                           /*   and the mask by rank; 1 the paths the default falls  */
                           /*   back to when their staging does not fit (per-entry   */
                           /*   appends, mask by degree); the counts are the same    */
    int32_t reserved[4];
} tc_schedule;

/* Fill *s with the defaults (autotuned table if one was loaded). */
int tc_schedule_default(tc_schedule *s);

typedef struct {
    double ms_total;        /* hipEvent time of the whole count               */
    double ms_prepare;      /* orientation + binning                          */
    double ms_count;        /* intersection kernels                           */
    double ms_bin[4];       /* per bin (small, mid, large, direct)            */
    double bytes_alg_bin[4];/* algorithmic bytes 4*(L_u+L_v) per bin          */
    int64_t tasks_bin[4];   /* intersection tasks per bin                     */
    int64_t jobs_bin[4];
    int64_t strategy_tasks[3]; /* tasks run as merge / bsearch / hash         */
    uint64_t triangles;     /* T for simple graphs (global/6), else 0         */
    int32_t method_used;
    int32_t graph_replay;   /* 1: this count was a hipGraph replay (capture)  */
    double items_bin[3];    /* task-list entries the tasks of each bin stream */
    double stream_bytes_bin[3]; /* bytes each bin must read: 4 per streamed
                                   entry + 24 per task (task entry and vertex
                                   record) + 4 per owner-tile entry per job */
    /* largest bin, oriented method: its hash-strategy tasks run in a kernel of
       their own, one launch for bitmap tiles [0] and one for hash-table tiles
       [1]; their hipEvent times, tasks, streamed entries and bytes (the same
       unit as stream_bytes_bin) */
    double ms_long[2];
    int64_t tasks_long[2];
    double items_long[2];
    double stream_bytes_long[2];
    int64_t mergepath_tasks; /* tasks run as a wavefront-wide merge path     */
} tc_stats;

/* ------------------------------------------------------ static count -- */
/* callDeviceAllTriangles (include/static_triangle_counting/cct.hpp:15-17,
 * src/static_triangle_counting/cct.cu:272-277): per_vertex_dev (device
 * uint32[nv], may be NULL) is overwritten with the per-vertex counts;
 * global_host (may be NULL) receives the uint64 total.  sched and stats may
 * be NULL.  Synchronous with respect to the handle's stream. */
int tc_static_count(tc_graph *g, uint32_t *per_vertex_dev, uint64_t *global_host,
                    const tc_schedule *sched, tc_stats *stats);

/* callDeviceAllTrianglesCSR (tests/cct_csr.cu:279-285): counts on a raw
 * device CSR with the CSR kernel's convention (entries with dest < src are
 * skipped, sum = 3T).  d_off int64[nv+1], d_ind int32[nnz], out uint32[nv]. */
int tc_csr_count_device(int32_t nv, int64_t nnz, const int64_t *d_off, const int32_t *d_ind,
                        uint32_t *out_dev, uint64_t *global_host, void *hip_stream);
/* No counterpart in the reference: tc_csr_count_device for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same arguments, output and
 * errors; the caller's CSR is read in place, and the only scratch is 256
 * bytes of counters in a static device array per GPU (validation flags, a
 * work cursor, the global sum), so no graph is too large for it.  Calls no
 * allocating routine.  Calls that overlap on one GPU share that scratch and
 * race; serialise them (one stream, or external locking). */
int tc_csr_count_device_noalloc(int32_t nv, int64_t nnz, const int64_t *d_off,
                                const int32_t *d_ind, uint32_t *out_dev, uint64_t *global_host,
                                void *hip_stream);

/* The same with the reference's int32 offsets (const length_t *d_off,
 * tests/cct_csr.cu:279-285): d_off int32[nv+1] on the device, widened there;
 * nnz = d_off[nv]. */
int tc_csr_count_device32(int32_t nv, const int32_t *d_off, const int32_t *d_ind, uint32_t *out_dev,
                          uint64_t *global_host, void *hip_stream);

/* ------------------------------------------------------ batch update -- */
typedef struct {
    int64_t inserted;        /* undirected pairs newly inserted                  */
    int64_t deleted;         /* undirected pairs removed                          */
    int64_t dup_inserts;     /* no-op inserts (in-batch repeat or present)        */
    int64_t missing_deletes; /* no-op deletes (absent pair or in-batch repeat)    */
    int64_t reallocations;   /* blocks moved to grow                              */
    double ms_update;        /* store update time                                 */
    double ms_delta;         /* delta counting time (0 in a replay: not timed)     */
    int32_t replayed;        /* 1: the batch ran as a recorded hipGraph            */
    int32_t reserved[7];
} tc_batch_stats;

/* Streaming batch: cuStinger edgeInsertionsSorted / edgeDeletions
 * (tests/cct_upd.cpp:395-398, tests/katz-tester.cu:244) fused with
 * callDeviceNewTriangles (src/streaming_triangle_counting/cct_upd.cu:412-440).
 * u, v: B undirected pairs (device int32); is_delete: device uint8[B] or NULL
 * (all insertions).  Deletions are counted on the pre-deletion graph and
 * removed first, then insertions are applied and counted.  Self-loops and
 * out-of-range ids are rejected (TC_ERR_INVALID, store untouched).
 * per_vertex_delta_dev (device uint32[nv], may be NULL) ACCUMULATES the
 * per-vertex delta (mod 2^32), like the reference's outPutTriangles;
 * delta_global_host (may be NULL) receives the signed change of the global
 * count (6 * change in T). */
int tc_batch_apply(tc_graph *g, const int32_t *u, const int32_t *v, const uint8_t *is_delete,
                   int64_t B, uint32_t *per_vertex_delta_dev, int64_t *delta_global_host,
                   const tc_schedule *sched, tc_batch_stats *stats);
/* With sched->capture, the batch is padded to its size class (256, then
 * steps of an eighth of the next power of two); the second batch of a class
 * (same output and buffers) records the sequence after validation as a
 * hipGraph and later batches of the class replay it.  The host then reads
 * back only twice per batch: the validation and the results. */

/* Store-only update (no counting): cuStinger edgeInsertions / edgeDeletions
 * of undirected pairs (both directions of each pair) on a symmetric store. */
int tc_batch_update_only(tc_graph *g, const int32_t *u, const int32_t *v,
                         const uint8_t *is_delete, int64_t B, tc_batch_stats *stats);

/* cuStinger edgeInsertions / edgeDeletions of DIRECTED entries, as
 * tests/katz-tester.cu:211-252 applies them to a directed store: each entry
 * (src[i], dst[i]) (device int32[B]) is inserted into, or with is_delete[i]
 * (device uint8[B], or NULL: all insertions) removed from, the row of src[i]
 * only.  Any store (symmetric or not, self-loops allowed); in-batch repeats,
 * present insertions and absent deletions are no-ops counted in the stats.
 * Deletions are applied before insertions.  A symmetric store stays flagged
 * symmetric when every touched entry's reverse ends in the same state;
 * otherwise symmetry is re-checked when next needed (tc_batch_apply refuses
 * an asymmetric store; tc_static_count counts it with the direct method). */
int tc_batch_update_directed(tc_graph *g, const int32_t *src, const int32_t *dst,
                             const uint8_t *is_delete, int64_t B, tc_batch_stats *stats);

/* BatchUpdate::sortDeviceBUD (tests/cct_upd.cpp:395) with the duplicate
 * flags the delta kernels consume (cct_upd.cu:252,308,364): the nb directed
 * device entries (src, dst) are sorted by (src, dst) into s_src/s_dst;
 * dup[i] = 1 for in-batch repeats (all but the first) and for entries already
 * in the store; boff (device int64[nv+1]) holds each vertex's first sorted
 * entry.  Call before inserting the batch. */
int tc_batch_prepare(tc_graph *g, int64_t nb, const int32_t *src, const int32_t *dst,
                     int32_t *s_src, int32_t *s_dst, int32_t *dup, int64_t *boff);

/* Reference-faithful delta kernels on a prepared batch, for parity of the
 * three passes (cct_upd.cu:231-410): s_src/s_dst sorted directed entries,
 * dup flags (device int32[nb]), boff int64[nv+1] batch offsets; the graph
 * must already contain the batch.  pass_mask bit0/1/2 selects the passes;
 * out_dev accumulates. */
int tc_stream_delta_passes(tc_graph *g, int64_t nb, const int32_t *s_src, const int32_t *s_dst,
                           const int32_t *dup, const int64_t *boff, uint32_t *out_dev,
                           int32_t pass_mask);

/* ---------------------------------------------------- batched graphs -- */
/* Batched static TC over G independent graphs concatenated into one CSR with
 * global vertex ids: graph i owns vertices [voff[i], voff[i+1]); its entries
 * must stay inside that range.  off int64[voff[G]+1], adj int32, all device
 * pointers.  per_graph_dev: device uint64[G] (global count per graph);
 * per_vertex_dev: device uint32[voff[G]] or NULL. */
int tc_batched_count(int32_t G, const int64_t *voff_dev, const int64_t *off_dev,
                     const int32_t *adj_dev, int64_t nnz, int32_t nv_total,
                     uint64_t *per_graph_dev, uint32_t *per_vertex_dev,
                     const tc_schedule *sched, tc_stats *stats, void *hip_stream);

/* The same input as a read-only handle, validated once and counted any number
 * of times (no counterpart in the reference: for callers that count the same
 * batch of graphs repeatedly).  The device arrays must outlive the handle,
 * which is released with tc_graph_destroy and cannot take batch updates. */
int tc_batched_graph_create(int32_t G, const int64_t *voff_dev, const int64_t *off_dev,
                            const int32_t *adj_dev, int64_t nnz, int32_t nv_total, void *hip_stream,
                            tc_graph **out);
int tc_batched_graph_count(tc_graph *g, uint64_t *per_graph_dev, uint32_t *per_vertex_dev,
                           const tc_schedule *sched, tc_stats *stats);

/* ----------------------------------------------------- load balancing -- */
/* cusLoadBalance::LoadBalanceArray (include/load_balance.cuh:77-102):
 * work[i] = degree of verts[i] (device int32 ids, len entries); returns the
 * partition count in *P and the lower-bound partition points in parts_dev
 * (device int64, room for 1000). */
int tc_lb_partition(tc_graph *g, const int32_t *verts_dev, int64_t len, int64_t *parts_dev,
                    int64_t *P);
/* Degree binning of the LB traversal (include/operators.cuh:164-288):
 * counts of vertices with degree >= 64, in [16,64) and < 16 among verts. */
int tc_lb_degree_bins(tc_graph *g, const int32_t *verts_dev, int64_t len, int64_t counts[3]);
/* Edge numbering of the load-balanced traversal (allVinA_TraverseEdges_LB,
 * include/operators.cuh:291-310, on the prefix of cusLoadBalance's work
 * array, src/load_balance.cu:32-54): pre_dev[i] = sum of the degrees of
 * verts_dev[0..i) for i in [0, len] (device int64, len+1 entries).  verts_dev
 * NULL means every vertex (len is then nv).  Runs on the handle's stream. */
int tc_lb_edge_prefix(tc_graph *g, const int32_t *verts_dev, int64_t len, int64_t *pre_dev);

/* ------------------------------------------------------ operator views -- */
/* Device view of the store for user kernels: what the reference's operators
 * read through cuStinger::devicePtr(), dVD->used[v] and dVD->adj[v]->dst
 * (include/operators.cuh:73-74, 137-140, 226-227).  Vertex v's neighbours are
 * pool[boff[v] .. boff[v]+len[v]), sorted ascending.  The pointers stay valid
 * until the next update or destroy; stream is the handle's stream, on which
 * kernels over the view are ordered after the library's own work.  Not
 * available on a batched handle.  include/custinger_ops.hpp builds the
 * traversal operators on it. */
typedef struct {
    int32_t nv;
    int32_t reserved;
    int64_t nnz;
    const int64_t *boff;
    const int32_t *len;
    const int32_t *pool;
    void *stream;
} tc_graph_view;
int tc_graph_get_view(tc_graph *g, tc_graph_view *view);

/* --------------------------------------------------------- graph files -- */
/* readGraphDIMACS / readGraphSNAP / readGraphMatrixMarket, the external
 * cuStinger utilities the reference drivers call (tests/cct.cpp:239-254,
 * tests/cct_upd.cpp:210-219): DIMACS/METIS ".graph" adjacency lists (1-based),
 * SNAP ".txt" edge lists (0-based, '#' comments) and MatrixMarket coordinate
 * ".mtx" files (1-based; "symmetric" stores one triangle).  TC_FMT_AUTO picks
 * the format from the name as the drivers do.  The result is a host CSR,
 * symmetrised, without self-loops, deduplicated and row-sorted unless
 * TC_READ_AS_IS (pairs as given, only sorted and deduplicated).  *off
 * (int64[nv+1]) and *adj (int32[nnz]) are malloc'd by the library; release
 * them with tc_host_free (or free). */
enum { TC_FMT_AUTO = 0, TC_FMT_DIMACS = 1, TC_FMT_SNAP = 2, TC_FMT_MTX = 3 };
enum { TC_READ_AS_IS = 0x1 };
int tc_read_graph(const char *path, int format, uint32_t flags, int32_t *nv, int64_t *nnz,
                  int64_t **off, int32_t **adj);
int tc_host_free(void *p);

/* ---------------------------------------------------------- generators -- */
/* Counter-based R-MAT / uniform pair generators and CSR canonicaliser
 * (symmetrise, drop self-loops, dedup, sort) on the device.  The same
 * generator is restated on the CPU so a graph can be rebuilt anywhere. */
int tc_gen_rmat_pairs(int32_t scale, int64_t m, double a, double b, double c, uint64_t seed,
                      int32_t *u_dev, int32_t *v_dev, void *hip_stream);
int tc_gen_er_pairs(int32_t n, int64_t m, uint64_t seed, int32_t *u_dev, int32_t *v_dev,
                    void *hip_stream);
/* Batched R-MAT: graph i (of G) has n[i] vertices, ef[i]*n[i] pairs drawn on
 * 2^ceil(log2 n[i]) with seed seeds[i]; pairs with an endpoint >= n[i] are
 * rejected (written as a self-loop, which canonicalisation drops).  Ids are
 * global (offset by the sum of the previous n).  Host arrays n, ef, seeds;
 * device outputs u, v hold sum(ef[i]*n[i]) pairs. */
int tc_gen_rmat_multi(int32_t G, const int32_t *n, const int32_t *ef, const uint64_t *seeds,
                      double a, double b, double c, int32_t *u_dev, int32_t *v_dev,
                      void *hip_stream);
/* Builds a canonical CSR from m device pairs; *off_dev (int64[nv+1]) and
 * *adj_dev are allocated by the call (free with tc_device_free).  Pairs with
 * an endpoint outside [0,nv) are dropped like self-loops. */
int tc_canon_csr_device(int32_t nv, int64_t m, const int32_t *u_dev, const int32_t *v_dev,
                        int64_t **off_dev, int32_t **adj_dev, int64_t *nnz, void *hip_stream);

/* ------------------------------------------------ learned cost model -- */
/* 3-layer MLP (4 -> 32 -> 32 -> 3) over (log2 |short|, log2 |long|,
 * log2 ratio, bin) predicting per-strategy cost; trained on the GPU. */
#define TC_MLP_IN 4
#define TC_MLP_HIDDEN 32
#define TC_MLP_OUT 3
#define TC_MLP_PARAMS (TC_MLP_IN * TC_MLP_HIDDEN + TC_MLP_HIDDEN + \
                       TC_MLP_HIDDEN * TC_MLP_HIDDEN + TC_MLP_HIDDEN + \
                       TC_MLP_HIDDEN * TC_MLP_OUT + TC_MLP_OUT)

/* Timing collection: runs every strategy on a sample of up to max_samples
 * intersection tasks of the graph and writes per-sample features
 * (float[4]) and measured cycles per strategy (float[3]) to host arrays.
 * Returns the number of samples in *n. */
int tc_collect_timings(tc_graph *g, int64_t max_samples, uint64_t seed, float *feat_host,
                       float *cost_host, int64_t *n);

/* One FP32 training step on device: forward, MSE loss on log-cost,
 * backward, SGD/Adam update of params_dev (float[TC_MLP_PARAMS]).
 * grads_dev (may be NULL) receives the gradient; loss_host the loss. */
int tc_mlp_train_step(float *params_dev, float *adam_m_dev, float *adam_v_dev, int32_t step,
                      const float *feat_dev, const float *target_dev, int64_t n, float lr,
                      float *grads_dev, float *loss_host);
/* FP32 forward on device (predictions float[n*3]). */
int tc_mlp_forward(const float *params_dev, const float *feat_dev, int64_t n, float *pred_dev);
/* BF16 inference over the quantised degree grid: fills the strategy table
 * used by TC_POLICY_LEARNED. */
int tc_model_load(const float *params_host);
int tc_model_table(uint8_t *table_host, int32_t *dim);
/* A 32 x 32 strategy table given directly (row: floor(log2 |owner tile|)+1,
 * column: floor(log2 |task list|)+1; entries 0 merge, 1 bsearch, 2 hash,
 * 3 merge path), used by TC_POLICY_LEARNED in place of a model's table
 * (the in-situ sweep's probes set one per run).  NULL unloads it. */
int tc_strategy_table_set(const uint8_t *table_host, int32_t dim);

/* -------------------------------------------------------- autotuning -- */
/* Sweeps bin thresholds and job grain on g and stores the best schedule as
 * the process default (the reference's exhaustive sweep, tests/cct.cpp:177-209). */
int tc_autotune(tc_graph *g, tc_schedule *best, double *best_ms);
/* Makes s the process default schedule (what tc_schedule_default returns and
 * what a NULL schedule means), e.g. one saved from an earlier tc_autotune;
 * NULL restores the built-in default.  No counterpart in the reference. */
int tc_schedule_set_default(const tc_schedule *s);

#ifdef __cplusplus
}
#endif
#endif /* CUSTINGER_TC_H */
