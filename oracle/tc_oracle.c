/*
 * tc_oracle.c -- CPU oracle for the triangle-counting path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product).  See tc_oracle.h
 * for the contract of each function and the reference lines it restates.
 * Pinning: tests/test_oracle.py checks every function here against closed
 * forms (K_n, wheels, friendship graphs, stars, cycles, bipartite graphs),
 * brute force on tiny graphs and the networkx/scipy fixtures under
 * tests/golden/.
 */
#include "tc_oracle.h"

#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------ RNG -- */

static inline uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

/* Counter-based: a keyed splitmix64 finaliser.  Identical formula in
 * custingeralg_amd/csrc/tc_gen.hip (tc_rng_dev). */
uint64_t tco_rng(uint64_t seed, uint64_t ctr)
{
    uint64_t key = mix64(seed + 0x9E3779B97F4A7C15ULL);
    return mix64(key ^ (ctr * 0x9E3779B97F4A7C15ULL + 0x632BE59BD9B4E019ULL));
}

static inline int32_t scale_to(uint64_t r, int64_t n)
{
    /* high 32 bits times n, shifted: uniform enough, exact integer math */
    return (int32_t)(((r >> 32) * (uint64_t)n) >> 32);
}

void tco_gen_er_pairs(int64_t n, int64_t m, uint64_t seed, int32_t *u, int32_t *v)
{
    for (int64_t i = 0; i < m; ++i) {
        u[i] = scale_to(tco_rng(seed, 2 * (uint64_t)i), n);
        v[i] = scale_to(tco_rng(seed, 2 * (uint64_t)i + 1), n);
    }
}

void tco_rmat_thresholds(double a, double b, double c, uint64_t t[3])
{
    const double s = 9007199254740992.0; /* 2^53 */
    t[0] = (uint64_t)(a * s);
    t[1] = (uint64_t)((a + b) * s);
    t[2] = (uint64_t)((a + b + c) * s);
}

void tco_gen_rmat_pairs(int scale, int64_t m, const uint64_t t[3], uint64_t seed,
                        int32_t *u, int32_t *v)
{
#pragma omp parallel for schedule(static) if (m > (1 << 20))
    for (int64_t i = 0; i < m; ++i) {
        uint32_t x = 0, y = 0;
        for (int l = 0; l < scale; ++l) {
            uint64_t r = tco_rng(seed, (uint64_t)i * 64 + (uint64_t)l) >> 11;
            uint32_t bit = 1u << (scale - 1 - l);
            if (r < t[0]) {
            } else if (r < t[1]) {
                y |= bit;
            } else if (r < t[2]) {
                x |= bit;
            } else {
                x |= bit;
                y |= bit;
            }
        }
        u[i] = (int32_t)x;
        v[i] = (int32_t)y;
    }
}

/* ------------------------------------------------------ canonical CSR -- */

static void radix_sort_u64(uint64_t *a, uint64_t *tmp, int64_t n)
{
    static int64_t cnt[65536];
    uint64_t *src = a, *dst = tmp;
    for (int pass = 0; pass < 4; ++pass) {
        int sh = pass * 16;
        memset(cnt, 0, sizeof(cnt));
        for (int64_t i = 0; i < n; ++i) cnt[(src[i] >> sh) & 0xFFFF]++;
        int trivial = 0;
        for (int d = 0; d < 65536; ++d)
            if (cnt[d] == n) trivial = 1;
        if (trivial) continue;
        int64_t s = 0;
        for (int d = 0; d < 65536; ++d) {
            int64_t c = cnt[d];
            cnt[d] = s;
            s += c;
        }
        for (int64_t i = 0; i < n; ++i) dst[cnt[(src[i] >> sh) & 0xFFFF]++] = src[i];
        uint64_t *t = src;
        src = dst;
        dst = t;
    }
    if (src != a) memcpy(a, src, (size_t)n * sizeof(uint64_t));
}

int64_t tco_canon_csr(int64_t n, int64_t m, const int32_t *u, const int32_t *v,
                      int64_t *off, int32_t *adj)
{
    uint64_t *k = (uint64_t *)malloc((size_t)(2 * m + 1) * sizeof(uint64_t));
    uint64_t *t = (uint64_t *)malloc((size_t)(2 * m + 1) * sizeof(uint64_t));
    int64_t nk = 0;
    for (int64_t i = 0; i < m; ++i) {
        /* self-loops and pairs with an endpoint outside [0,n) are dropped */
        if (u[i] == v[i] || u[i] < 0 || v[i] < 0 || u[i] >= n || v[i] >= n) continue;
        k[nk++] = ((uint64_t)(uint32_t)u[i] << 32) | (uint32_t)v[i];
        k[nk++] = ((uint64_t)(uint32_t)v[i] << 32) | (uint32_t)u[i];
    }
    radix_sort_u64(k, t, nk);
    int64_t nnz = 0;
    memset(off, 0, (size_t)(n + 1) * sizeof(int64_t));
    for (int64_t i = 0; i < nk; ++i) {
        if (i > 0 && k[i] == k[i - 1]) continue;
        int32_t s = (int32_t)(k[i] >> 32);
        adj[nnz++] = (int32_t)(k[i] & 0xFFFFFFFFu);
        off[s + 1]++;
    }
    for (int64_t i = 0; i < n; ++i) off[i + 1] += off[i];
    free(k);
    free(t);
    return nnz;
}

/* ------------------------------------------- reference CPU restatement -- */

/* tests/cct.cpp:32-69: merge with forward skip loops, early exit on empty
 * or disjoint ranges.  The self-edge skip block there is compiled out
 * (#if(0)), so it is not restated. */
int32_t tco_host_single_intersection(int32_t ai, int64_t alen, const int32_t *a,
                                     int32_t bi, int64_t blen, const int32_t *b)
{
    (void)ai;
    (void)bi;
    int64_t ka = 0, kb = 0;
    int32_t out = 0;
    if (!alen || !blen || a[alen - 1] < b[0] || b[blen - 1] < a[0]) return 0;
    while (1) {
        if (ka >= alen || kb >= blen) break;
        int32_t va = a[ka], vb = b[kb];
        if (va == vb) {
            ++ka;
            ++kb;
            ++out;
        } else if (va < vb) {
            ++ka;
            while (ka < alen && a[ka] < vb) ++ka;
        } else {
            ++kb;
            while (kb < blen && va > b[kb]) ++kb;
        }
    }
    return out;
}

/* tests/cct.cpp:71-89 over rows [v0,v1). */
static int64_t host_count_range(const int64_t *off, const int32_t *ind, int32_t *triNE,
                                int32_t v0, int32_t v1, int nthreads)
{
    int64_t sum = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : sum) num_threads(nthreads) if (nthreads > 1)
    for (int32_t src = v0; src < v1; src++) {
        int64_t srcLen = off[src + 1] - off[src];
        for (int64_t it = off[src]; it < off[src + 1]; it++) {
            int32_t dest = ind[it];
            int64_t destLen = off[dest + 1] - off[dest];
            int32_t c = tco_host_single_intersection(src, srcLen, ind + off[src], dest,
                                                     destLen, ind + off[dest]);
            if (triNE) triNE[it] = c;
            sum += c;
        }
    }
    return sum;
}

int64_t tco_host_count_triangles(int32_t nv, const int64_t *off, const int32_t *ind,
                                 int32_t *triNE, int nthreads)
{
    return host_count_range(off, ind, triNE, 0, nv, nthreads);
}

/* The per-entry body of hostCountTriangles (tests/cct.cpp:79-86) on a sample
 * of entry indices: the CPU-baseline leg times this on uniformly drawn
 * entries.  Returns the int64 sum of the sampled intersections. */
int64_t tco_host_count_entries(int32_t nv, const int64_t *off, const int32_t *ind,
                               const int64_t *entries, int64_t n, int nthreads)
{
    int64_t sum = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : sum) num_threads(nthreads) if (nthreads > 1)
    for (int64_t i = 0; i < n; ++i) {
        int64_t e = entries[i];
        int32_t lo = 0, hi = nv - 1; /* row holding entry e */
        while (lo < hi) {
            int32_t mid = lo + ((hi - lo + 1) >> 1);
            if (off[mid] <= e) lo = mid; else hi = mid - 1;
        }
        int32_t src = lo, dest = ind[e];
        sum += tco_host_single_intersection(src, off[src + 1] - off[src], ind + off[src], dest,
                                            off[dest + 1] - off[dest], ind + off[dest]);
    }
    return sum;
}

int64_t tco_host_count_rows(int32_t nv, const int64_t *off, const int32_t *ind,
                            int32_t v0, int32_t v1, int nthreads)
{
    if (v0 < 0) v0 = 0;
    if (v1 > nv) v1 = nv;
    return host_count_range(off, ind, NULL, v0, v1, nthreads);
}

/* Sequential merge count: the exact behaviour of intersectPath
 * (cct.cu:145-166) at one thread per intersection, where found=1, the start
 * point is (0,0) and the trailing correction never fires. */
static uint32_t merge_count(const int32_t *u, int64_t ul, const int32_t *v, int64_t vl)
{
    int64_t i = 0, j = 0;
    uint32_t c = 0;
    while (i < ul && j < vl) {
        int32_t cmp = u[i] - v[j];
        c += (cmp == 0);
        i += (cmp <= 0);
        j += (cmp >= 0);
    }
    return c;
}

static void custinger_like(int32_t nv, const int64_t *off, const int32_t *ind, uint32_t *out,
                           int skip_lower, int nthreads)
{
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads) if (nthreads > 1)
    for (int32_t src = 0; src < nv; ++src) {
        int64_t srcLen = off[src + 1] - off[src];
        uint32_t t = 0;
        for (int64_t k = off[src]; k < off[src + 1]; ++k) {
            int32_t dest = ind[k];
            if (skip_lower && dest < src) continue; /* tests/cct_csr.cu:252-253 */
            int64_t destLen = off[dest + 1] - off[dest];
            /* cct.cu:247-249 */
            if (src == dest || destLen < 2 || srcLen < 2) continue;
            /* cct.cu:251-255: shorter list first; the count is symmetric */
            t += merge_count(ind + off[src], srcLen, ind + off[dest], destLen);
        }
        out[src] = t;
    }
}

void tco_custinger_count(int32_t nv, const int64_t *off, const int32_t *ind, uint32_t *out,
                         int nthreads)
{
    custinger_like(nv, off, ind, out, 0, nthreads);
}

void tco_csr_count(int32_t nv, const int64_t *off, const int32_t *ind, uint32_t *out,
                   int nthreads)
{
    custinger_like(nv, off, ind, out, 1, nthreads);
}

int64_t tco_sum_triangle_array(const int32_t *h, int32_t nv)
{
    int64_t s = 0;
    for (int32_t i = 0; i < nv; ++i) s += h[i];
    return s;
}

/* ------------------------------------------------------ forward count -- */

uint64_t tco_forward_count(int32_t nv, const int64_t *off, const int32_t *ind, uint64_t *t,
                           int nthreads)
{
    /* orient a->b iff (deg a, a) < (deg b, b); N+ lists keep id order */
    int64_t *poff = (int64_t *)calloc((size_t)nv + 1, sizeof(int64_t));
#pragma omp parallel for num_threads(nthreads) if (nthreads > 1)
    for (int32_t a = 0; a < nv; ++a) {
        int64_t da = off[a + 1] - off[a], c = 0;
        for (int64_t k = off[a]; k < off[a + 1]; ++k) {
            int32_t b = ind[k];
            int64_t db = off[b + 1] - off[b];
            c += (da < db) || (da == db && a < b);
        }
        poff[a + 1] = c;
    }
    for (int32_t a = 0; a < nv; ++a) poff[a + 1] += poff[a];
    int32_t *pind = (int32_t *)malloc((size_t)(poff[nv] + 1) * sizeof(int32_t));
#pragma omp parallel for num_threads(nthreads) if (nthreads > 1)
    for (int32_t a = 0; a < nv; ++a) {
        int64_t da = off[a + 1] - off[a], p = poff[a];
        for (int64_t k = off[a]; k < off[a + 1]; ++k) {
            int32_t b = ind[k];
            int64_t db = off[b + 1] - off[b];
            if ((da < db) || (da == db && a < b)) pind[p++] = b;
        }
    }
    if (t) memset(t, 0, (size_t)nv * sizeof(uint64_t));
    uint64_t total = 0;
#pragma omp parallel num_threads(nthreads) if (nthreads > 1)
    {
        int32_t *stamp = (int32_t *)malloc((size_t)nv * sizeof(int32_t));
        for (int32_t i = 0; i < nv; ++i) stamp[i] = -1;
        uint64_t local = 0;
#pragma omp for schedule(dynamic, 64)
        for (int32_t a = 0; a < nv; ++a) {
            if (poff[a + 1] - poff[a] < 2) continue;
            for (int64_t k = poff[a]; k < poff[a + 1]; ++k) stamp[pind[k]] = a;
            uint64_t ta = 0;
            for (int64_t k = poff[a]; k < poff[a + 1]; ++k) {
                int32_t b = pind[k];
                uint64_t tb = 0;
                for (int64_t q = poff[b]; q < poff[b + 1]; ++q) {
                    int32_t w = pind[q];
                    if (stamp[w] == a) {
                        ++tb;
                        if (t) __atomic_fetch_add(&t[w], 1, __ATOMIC_RELAXED);
                    }
                }
                if (tb && t) __atomic_fetch_add(&t[b], tb, __ATOMIC_RELAXED);
                ta += tb;
            }
            if (ta && t) __atomic_fetch_add(&t[a], ta, __ATOMIC_RELAXED);
            local += ta;
        }
        free(stamp);
        __atomic_fetch_add(&total, local, __ATOMIC_RELAXED);
    }
    free(poff);
    free(pind);
    return total;
}

/* ---------------------------------------------------------- streaming -- */

static int has_edge(const int64_t *off, const int32_t *ind, int32_t s, int32_t d)
{
    int64_t lo = off[s], hi = off[s + 1];
    while (lo < hi) {
        int64_t mid = (lo + hi) >> 1;
        if (ind[mid] < d)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo < off[s + 1] && ind[lo] == d;
}

void tco_batch_prepare(int32_t nv, int64_t nb, const int32_t *src, const int32_t *dst,
                       const int64_t *g_off, const int32_t *g_ind, int32_t *s_src,
                       int32_t *s_dst, int32_t *dup, int64_t *boff)
{
    uint64_t *k = (uint64_t *)malloc((size_t)(nb + 1) * sizeof(uint64_t));
    uint64_t *t = (uint64_t *)malloc((size_t)(nb + 1) * sizeof(uint64_t));
    for (int64_t i = 0; i < nb; ++i)
        k[i] = ((uint64_t)(uint32_t)src[i] << 32) | (uint32_t)dst[i];
    radix_sort_u64(k, t, nb);
    memset(boff, 0, (size_t)(nv + 1) * sizeof(int64_t));
    for (int64_t i = 0; i < nb; ++i) {
        s_src[i] = (int32_t)(k[i] >> 32);
        s_dst[i] = (int32_t)(k[i] & 0xFFFFFFFFu);
        int rep = (i > 0 && k[i] == k[i - 1]);
        dup[i] = rep || has_edge(g_off, g_ind, s_src[i], s_dst[i]);
        boff[s_src[i] + 1]++;
    }
    for (int32_t i = 0; i < nv; ++i) boff[i + 1] += boff[i];
    free(k);
    free(t);
}

/* intersectCount<uMasked,vMasked,subtract,upd3rdV> at one thread per
 * intersection (cct_upd.cu:146-176): a masked entry advances its own side
 * without counting; every match may update the third vertex. */
static uint32_t masked_merge(const int32_t *u, int64_t ul, const int32_t *um, const int32_t *v,
                             int64_t vl, const int32_t *vm, uint32_t *out, int sign)
{
    int64_t i = 0, j = 0;
    uint32_t c = 0;
    while (i < ul && j < vl) {
        int vmask = vm ? vm[j] : 0;
        int umask = um ? um[i] : 0;
        int32_t cmp = u[i] - v[j];
        int hit = (cmp == 0 && !umask && !vmask);
        c += hit;
        if (hit && out) out[u[i]] += (uint32_t)sign;
        int64_t di = ((cmp <= 0 && !vmask) || umask);
        int64_t dj = ((cmp >= 0 && !umask) || vmask);
        i += di;
        j += dj;
    }
    return c;
}

void tco_stream_delta(int32_t nv, const int64_t *gp_off, const int32_t *gp_ind, int64_t nb,
                      const int32_t *s_src, const int32_t *s_dst, const int32_t *dup,
                      const int64_t *boff, uint32_t *out, int pass_mask)
{
    (void)nv;
    /* pass 1: devicecuStingerNewTriangles (:231-285) */
    if (pass_mask & 1)
        for (int64_t e = 0; e < nb; ++e) {
            if (dup[e] == 1) continue;
            int32_t s = s_src[e], d = s_dst[e];
            int64_t sl = gp_off[s + 1] - gp_off[s], dl = gp_off[d + 1] - gp_off[d];
            if (s == d || dl < 2 || sl < 2) continue;
            int ss = sl < dl;
            int32_t sm = ss ? s : d, lg = ss ? d : s;
            int64_t sml = ss ? sl : dl, lgl = ss ? dl : sl;
            uint32_t c = masked_merge(gp_ind + gp_off[sm], sml, NULL, gp_ind + gp_off[lg], lgl,
                                      NULL, out, +1);
            out[s] += c;
            out[d] += c;
        }
    /* pass 2: deviceBUThreeTriangles (:287-341) */
    if (pass_mask & 2)
        for (int64_t e = 0; e < nb; ++e) {
            if (dup[e]) continue;
            int32_t s = s_src[e], d = s_dst[e];
            int64_t sl = boff[s + 1] - boff[s], dl = boff[d + 1] - boff[d];
            if (s == d || dl < 2 || sl < 2) continue;
            int ss = sl < dl;
            int32_t sm = ss ? s : d, lg = ss ? d : s;
            int64_t sml = ss ? sl : dl, lgl = ss ? dl : sl;
            uint32_t c = masked_merge(s_dst + boff[sm], sml, dup + boff[sm], s_dst + boff[lg],
                                      lgl, dup + boff[lg], NULL, 0);
            out[s] += c;
        }
    /* pass 3: deviceBUTwoCUOneTriangles (:343-410) */
    if (pass_mask & 4)
        for (int64_t e = 0; e < nb; ++e) {
            if (dup[e]) continue;
            int32_t s = s_src[e], d = s_dst[e];
            int64_t sl = boff[s + 1] - boff[s], dl = gp_off[d + 1] - gp_off[d];
            if (s == d || dl < 2 || sl < 2) continue;
            const int32_t *sp = s_dst + boff[s], *smk = dup + boff[s];
            const int32_t *dp = gp_ind + gp_off[d];
            uint32_t c;
            if (sl < dl)
                c = masked_merge(sp, sl, smk, dp, dl, NULL, out, -1);
            else
                c = masked_merge(dp, dl, NULL, sp, sl, smk, out, -1);
            out[s] -= c;
            out[d] -= c;
        }
}

/* -------------------------------------------------------- load balance -- */

int64_t tco_lb_partition(int64_t len, const int64_t *work, int64_t *prefix, int64_t *parts)
{
    /* estimateWorkPerVertex + scan over len+1 entries (load_balance.cuh:79-81):
     * prefix[i] = sum of work[0..i) for i in [0,len] */
    int64_t s = 0;
    for (int64_t i = 0; i < len; ++i) {
        prefix[i] = s;
        s += work[i];
    }
    prefix[len] = s;
    /* :83-96 */
    float avg = (float)s / (float)len;
    int64_t P = avg < 4 ? 200 : (avg < 7 ? 500 : 1000);
    if (P > len) P = len;
    /* deviceCreateNeedles (load_balance.cu:57-61) */
    for (int64_t b = 0; b < P; ++b) {
        int64_t needle = (int64_t)((double)prefix[len] * ((double)b / (double)P));
        /* sorted_search<bounds_lower> over prefix[0..len-1) (load_balance.cuh:100-101) */
        int64_t lo = 0, hi = len - 1;
        while (lo < hi) {
            int64_t mid = (lo + hi) >> 1;
            if (prefix[mid] < needle)
                lo = mid + 1;
            else
                hi = mid;
        }
        parts[b] = lo;
    }
    return P;
}
