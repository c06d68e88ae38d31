/*
 * tc_oracle.h -- CPU oracle for the triangle-counting path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product (custingeralg_amd/, libcustinger_tc.so) never links or calls it.
 *
 * Every function restates a reference routine (file:line relative to the
 * cuStingerAlg tree) or is a textbook routine pinned by closed forms and
 * brute force in tests/test_oracle.py.
 *
 * Conventions
 *   - vertex ids: int32, offsets: int64, CSR rows sorted ascending.
 *   - per-vertex counts: uint32 (same bits as the reference's int32
 *     triangle_t, include/static_triangle_counting/cct.hpp:12).
 */
#ifndef TC_ORACLE_H
#define TC_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- counter-based RNG shared bit-for-bit with the GPU generator ---- */
uint64_t tco_rng(uint64_t seed, uint64_t ctr);

/* Uniform pairs: pair i -> (u,v), u = rng(seed,2i) scaled to [0,n), v = rng(seed,2i+1). */
void tco_gen_er_pairs(int64_t n, int64_t m, uint64_t seed, int32_t *u, int32_t *v);

/* R-MAT pairs without noise.  Quadrant thresholds are 53-bit integers
 * t1=floor(a*2^53), t2=floor((a+b)*2^53), t3=floor((a+b+c)*2^53); level l of
 * pair i draws r = rng(seed, i*64+l) >> 11.  Recursive quadrant choice follows
 * the driver's rmat_edge (tests/cct_upd.cpp:67-112) minus its +-5% noise. */
void tco_rmat_thresholds(double a, double b, double c, uint64_t t[3]);
void tco_gen_rmat_pairs(int scale, int64_t m, const uint64_t t[3], uint64_t seed,
                        int32_t *u, int32_t *v);

/* Canonicalise undirected pairs into a CSR: symmetrise, drop self-loops,
 * dedup, sort ascending.  off has n+1 entries.  Returns nnz; adj must hold
 * 2*m entries (upper bound). */
int64_t tco_canon_csr(int64_t n, int64_t m, const int32_t *u, const int32_t *v,
                      int64_t *off, int32_t *adj);

/* ---- restatements of the reference counting semantics ---- */

/* hostSingleIntersection (tests/cct.cpp:32-69). */
int32_t tco_host_single_intersection(int32_t ai, int64_t alen, const int32_t *a,
                                     int32_t bi, int64_t blen, const int32_t *b);
/* hostCountTriangles (tests/cct.cpp:71-89): per-entry counts + int64 total.
 * triNE may be NULL.  nthreads>1 runs rows in parallel (the CPU-ref-NT leg). */
int64_t tco_host_count_triangles(int32_t nv, const int64_t *off, const int32_t *ind,
                                 int32_t *triNE, int nthreads);
/* Rows [v0,v1) only, for bounded CPU-baseline samples.  Returns the partial
 * int64 sum over those rows. */
int64_t tco_host_count_rows(int32_t nv, const int64_t *off, const int32_t *ind,
                            int32_t v0, int32_t v1, int nthreads);

/* The per-entry body of hostCountTriangles on sampled entry indices. */
int64_t tco_host_count_entries(int32_t nv, const int64_t *off, const int32_t *ind,
                               const int64_t *entries, int64_t n, int nthreads);
/* devicecuStingerAllTriangles semantics (src/static_triangle_counting/cct.cu:216-270)
 * at its exact (tpi=1) behaviour: out[src] = sum over entries dest of src with
 * src!=dest, deg(dest)>=2, deg(src)>=2 of |N(src) cap N(dest)|, mod 2^32. */
void tco_custinger_count(int32_t nv, const int64_t *off, const int32_t *ind,
                         uint32_t *out, int nthreads);
/* count_all_trianglesGPU semantics (tests/cct_csr.cu:228-276): as above but
 * entries with dest<src are skipped (sum = 3T). */
void tco_csr_count(int32_t nv, const int64_t *off, const int32_t *ind,
                   uint32_t *out, int nthreads);
/* sumTriangleArray (tests/cct.cpp:119-125): int64 sum of sign-extended int32. */
int64_t tco_sum_triangle_array(const int32_t *h, int32_t nv);

/* Degree-oriented forward count on a symmetric simple CSR: returns T and fills
 * t[v] = triangles at v (uint64).  t may be NULL. */
uint64_t tco_forward_count(int32_t nv, const int64_t *off, const int32_t *ind,
                           uint64_t *t, int nthreads);

/* ---- streaming (src/streaming_triangle_counting/cct_upd.cu) ---- */

/* Batch preparation with the semantics the driver relies on
 * (tests/cct_upd.cpp:313-332, 395): sort directed entries by (src,dst),
 * flag in-batch repeats (all but the first) and entries already present in
 * the pre-insert graph G; build batch offsets boff[nv+1] over the sorted batch
 * (flagged entries included).  Outputs have nb entries. */
void tco_batch_prepare(int32_t nv, int64_t nb, const int32_t *src, const int32_t *dst,
                       const int64_t *g_off, const int32_t *g_ind,
                       int32_t *s_src, int32_t *s_dst, int32_t *dup, int64_t *boff);

/* The three delta kernels at tpi=1, applied to the post-insert graph G' and
 * the prepared batch; out accumulates (uint32 wrap).  pass_mask selects
 * passes (bit0 devicecuStingerNewTriangles :231-285, bit1 deviceBUThreeTriangles
 * :287-341, bit2 deviceBUTwoCUOneTriangles :343-410). */
void tco_stream_delta(int32_t nv, const int64_t *gp_off, const int32_t *gp_ind,
                      int64_t nb, const int32_t *s_src, const int32_t *s_dst,
                      const int32_t *dup, const int64_t *boff, uint32_t *out,
                      int pass_mask);

/* ---- load balance (include/load_balance.cuh:77-102, src/load_balance.cu:32-66) ---- */
/* Returns the partition count P and fills parts[0..P) with lower-bound
 * partition points of needles total*b/P in the exclusive prefix of work. */
int64_t tco_lb_partition(int64_t len, const int64_t *work, int64_t *prefix,
                         int64_t *parts);

#ifdef __cplusplus
}
#endif
#endif
