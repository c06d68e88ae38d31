// tc_gen.hip -- synthetic graph generation on the device.
//
// Counter-based generators (bit-identical to the CPU restatement in the
// oracle): pair i of an R-MAT draw uses one 64-bit number per level,
// r = rng(seed, 64*i + level) >> 11, compared with 53-bit integer quadrant
// thresholds -- the recursive quadrant choice of rmat_edge
// (tests/cct_upd.cpp:67-112) without its per-level noise.  The canonicaliser
// symmetrises, drops self-loops, deduplicates and sorts with our own radix
// sort, then builds int64 row offsets.
#include "tc_internal.hpp"

namespace tc {

namespace {

__device__ __host__ inline uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

__device__ inline uint64_t rng_keyed(uint64_t key, uint64_t ctr)
{
    return mix64(key ^ (ctr * 0x9E3779B97F4A7C15ULL + 0x632BE59BD9B4E019ULL));
}

__device__ inline int32_t scale_to(uint64_t r, int64_t n)
{
    return (int32_t)(((r >> 32) * (uint64_t)n) >> 32);
}

__global__ void k_gen_rmat(int32_t scale, int64_t m, uint64_t t0, uint64_t t1, uint64_t t2,
                           uint64_t key, int32_t *__restrict__ u, int32_t *__restrict__ v)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m;
         i += (int64_t)gridDim.x * blockDim.x) {
        uint32_t x = 0, y = 0;
        for (int l = 0; l < scale; ++l) {
            uint64_t r = rng_keyed(key, (uint64_t)i * 64 + (uint64_t)l) >> 11;
            uint32_t bit = 1u << (scale - 1 - l);
            uint32_t qx = (r >= t1) ? 1u : 0u;                 // quadrants 3,4 set the row bit
            uint32_t qy = (r >= t0 && r < t1) || (r >= t2);    // quadrants 2,4 set the column bit
            x |= qx ? bit : 0u;
            y |= qy ? bit : 0u;
        }
        u[i] = (int32_t)x;
        v[i] = (int32_t)y;
    }
}

__global__ void k_gen_er(int32_t n, int64_t m, uint64_t key, int32_t *__restrict__ u,
                         int32_t *__restrict__ v)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m;
         i += (int64_t)gridDim.x * blockDim.x) {
        u[i] = scale_to(rng_keyed(key, 2 * (uint64_t)i), n);
        v[i] = scale_to(rng_keyed(key, 2 * (uint64_t)i + 1), n);
    }
}

// Both directions of every non-loop pair; loops become the all-ones key,
// which sorts after every valid key.
__global__ void k_pair_keys(const int32_t *__restrict__ u, const int32_t *__restrict__ v, int64_t m,
                            int32_t nv, int bits_v, uint64_t invalid, uint64_t *__restrict__ keys)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m;
         i += (int64_t)gridDim.x * blockDim.x) {
        uint64_t a = (uint32_t)u[i], b = (uint32_t)v[i];
        // self-loops and pairs with an endpoint outside [0,nv) are dropped
        bool loop = a == b || a >= (uint64_t)nv || b >= (uint64_t)nv;
        keys[2 * i] = loop ? invalid : ((a << bits_v) | b);
        keys[2 * i + 1] = loop ? invalid : ((b << bits_v) | a);
    }
}

__global__ void k_unique_flags(const uint64_t *__restrict__ keys, int64_t n, uint64_t invalid,
                               int32_t *__restrict__ flag)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        uint64_t k = keys[i];
        flag[i] = (k != invalid) && (i == 0 || keys[i - 1] != k);
    }
}

// Compacts unique keys into adj and writes row starts for the rows in
// (previous row, this row]; rows after the last key get nnz.
__global__ void k_compact_csr(const uint64_t *__restrict__ keys, const int32_t *__restrict__ flag,
                              const int64_t *__restrict__ pos, int64_t n, int bits_v,
                              int32_t nv, int32_t *__restrict__ adj, int64_t *__restrict__ off)
{
    const uint64_t vmask = (1ull << bits_v) - 1;
    const int64_t nnz = pos[n];
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        if (!flag[i]) continue;
        uint64_t k = keys[i];
        int32_t r = (int32_t)(k >> bits_v);
        int64_t p = pos[i];
        adj[p] = (int32_t)(k & vmask);
        // first entry of its row?  (the previous unique entry has a smaller row)
        int32_t rprev = -1;
        if (p > 0) {
            // the previous unique key is the last flagged key before i
            int64_t j = i - 1;
            while (!flag[j]) --j;
            rprev = (int32_t)(keys[j] >> bits_v);
        }
        if (rprev != r)
            for (int32_t q = rprev + 1; q <= r; ++q) off[q] = p;
        // last unique entry fills the tail
        if (p == nnz - 1)
            for (int32_t q = r + 1; q <= nv; ++q) off[q] = nnz;
    }
}

__global__ void k_fill_off(int64_t *off, int32_t nv)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i <= nv;
         i += (int64_t)gridDim.x * blockDim.x)
        off[i] = 0;
}

uint64_t host_key(uint64_t seed) { return mix64(seed + 0x9E3779B97F4A7C15ULL); }

struct MultiGraph {
    int32_t scale;
    int32_t n;
    int64_t voff;   // first global vertex id
    int64_t poff;   // first pair index
    uint64_t key;
};

// Pair p belongs to the graph whose pair range holds it; rejected pairs
// (an endpoint >= n) become the self-loop (voff, voff).
__global__ void k_gen_rmat_multi(const MultiGraph *__restrict__ gs, int32_t G, int64_t m,
                                 uint64_t t0, uint64_t t1, uint64_t t2, int32_t *__restrict__ u,
                                 int32_t *__restrict__ v)
{
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < m;
         p += (int64_t)gridDim.x * blockDim.x) {
        int32_t lo = 0, hi = G - 1;
        while (lo < hi) {
            const int32_t mid = (lo + hi + 1) >> 1;
            if (gs[mid].poff <= p) lo = mid; else hi = mid - 1;
        }
        const MultiGraph g = gs[lo];
        const uint64_t i = (uint64_t)(p - g.poff);
        uint32_t x = 0, y = 0;
        for (int l = 0; l < g.scale; ++l) {
            const uint64_t r = rng_keyed(g.key, i * 64 + (uint64_t)l) >> 11;
            const uint32_t bit = 1u << (g.scale - 1 - l);
            x |= (r >= t1) ? bit : 0u;
            y |= ((r >= t0 && r < t1) || (r >= t2)) ? bit : 0u;
        }
        const bool ok = (int32_t)x < g.n && (int32_t)y < g.n;
        u[p] = (int32_t)(g.voff + (ok ? x : 0));
        v[p] = (int32_t)(g.voff + (ok ? y : 0));
    }
}

}  // namespace

int gen_rmat(int32_t scale, int64_t m, double a, double b, double c, uint64_t seed, int32_t *u,
             int32_t *v, hipStream_t s)
{
    if (m == 0) return TC_OK;
    const double two53 = 9007199254740992.0;
    uint64_t t0 = (uint64_t)(a * two53), t1 = (uint64_t)((a + b) * two53),
             t2 = (uint64_t)((a + b + c) * two53);
    hipLaunchKernelGGL(k_gen_rmat, dim3(grid_for(m, 256, 1 << 16)), dim3(256), 0, s, scale, m, t0,
                       t1, t2, host_key(seed), u, v);
    TC_HIP(hipGetLastError());
    TC_HIP(hipStreamSynchronize(s));
    return TC_OK;
}

int gen_rmat_multi(int32_t G, const int32_t *n, const int32_t *ef, const uint64_t *seeds, double a,
                   double b, double c, int32_t *u, int32_t *v, hipStream_t s)
{
    std::vector<MultiGraph> gs((size_t)G);
    int64_t voff = 0, poff = 0;
    for (int32_t i = 0; i < G; ++i) {
        if (n[i] < 1 || ef[i] < 0) {
            set_last_error(TC_ERR_INVALID, "tc_gen_rmat_multi: n must be >= 1 and ef >= 0");
            return TC_ERR_INVALID;
        }
        int sc = 0;
        while ((1 << sc) < n[i]) ++sc;
        gs[i] = MultiGraph{sc < 1 ? 1 : sc, n[i], voff, poff, host_key(seeds[i])};
        voff += n[i];
        poff += (int64_t)ef[i] * n[i];
    }
    if (voff > INT32_MAX) {
        set_last_error(TC_ERR_CAPACITY, "tc_gen_rmat_multi: more than 2^31-1 vertices");
        return TC_ERR_CAPACITY;
    }
    if (poff == 0) return TC_OK;
    const double two53 = 9007199254740992.0;
    uint64_t t0 = (uint64_t)(a * two53), t1 = (uint64_t)((a + b) * two53),
             t2 = (uint64_t)((a + b + c) * two53);
    DevBuf dg;
    TC_TRY(dg.ensure(gs.size() * sizeof(MultiGraph)));
    TC_HIP(hipMemcpyAsync(dg.p, gs.data(), gs.size() * sizeof(MultiGraph), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_gen_rmat_multi, dim3(grid_for(poff, 256, 1 << 16)), dim3(256), 0, s,
                       dg.as<MultiGraph>(), G, poff, t0, t1, t2, u, v);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    dg.release();
    if (e != hipSuccess) {
        set_last_error(TC_ERR_HIP, std::string("tc_gen_rmat_multi: ") + hipGetErrorString(e));
        return TC_ERR_HIP;
    }
    return TC_OK;
}

int gen_er(int32_t n, int64_t m, uint64_t seed, int32_t *u, int32_t *v, hipStream_t s)
{
    if (m == 0) return TC_OK;
    hipLaunchKernelGGL(k_gen_er, dim3(grid_for(m, 256, 1 << 16)), dim3(256), 0, s, n, m,
                       host_key(seed), u, v);
    TC_HIP(hipGetLastError());
    TC_HIP(hipStreamSynchronize(s));
    return TC_OK;
}

int canon_csr_device(int32_t nv, int64_t m, const int32_t *u, const int32_t *v, int64_t **off_out,
                     int32_t **adj_out, int64_t *nnz_out, hipStream_t s)
{
    int bits_v = 1;
    while ((int64_t(1) << bits_v) < nv) ++bits_v;
    const int bits = 2 * bits_v;
    const uint64_t invalid = (bits >= 64) ? ~0ull : ((1ull << bits) - 1);
    const int64_t n = 2 * m;
    DevBuf keys, alt_keys, alt_vals, tmp, flag, pos;
    int64_t *off = nullptr;
    int32_t *adj = nullptr;
    int64_t nnz = 0;
    auto cleanup = [&]() {
        keys.release();
        alt_keys.release();
        alt_vals.release();
        tmp.release();
        flag.release();
        pos.release();
    };
    int st = TC_OK;
    if (hipMalloc(&off, (size_t)(nv + 1) * sizeof(int64_t)) != hipSuccess) {
        set_last_error(TC_ERR_OOM, "canon: offsets");
        return TC_ERR_OOM;
    }
    hipLaunchKernelGGL(k_fill_off, dim3(grid_for(nv + 1, 256, 65536)), dim3(256), 0, s, off, nv);
    if (n > 0) {
        if ((st = keys.ensure((size_t)n * sizeof(uint64_t))) != TC_OK) goto fail;
        hipLaunchKernelGGL(k_pair_keys, dim3(grid_for(m, 256, 1 << 16)), dim3(256), 0, s, u, v, m,
                           nv, bits_v, invalid, keys.as<uint64_t>());
        if ((st = radix_sort_u64(keys.as<uint64_t>(), nullptr, n, bits, alt_keys, alt_vals, tmp, s)) !=
            TC_OK)
            goto fail;
        alt_keys.release();
        if ((st = flag.ensure((size_t)n * sizeof(int32_t))) != TC_OK) goto fail;
        if ((st = pos.ensure((size_t)(n + 1) * sizeof(int64_t))) != TC_OK) goto fail;
        hipLaunchKernelGGL(k_unique_flags, dim3(grid_for(n, 256, 1 << 16)), dim3(256), 0, s,
                           keys.as<uint64_t>(), n, invalid, flag.as<int32_t>());
        if ((st = scan_counts_i32(flag.as<int32_t>(), pos.as<int64_t>(), n, tmp, s)) != TC_OK) goto fail;
        if (hipMemcpyAsync(&nnz, pos.as<int64_t>() + n, sizeof(int64_t), hipMemcpyDeviceToHost, s) !=
                hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess) {
            st = TC_ERR_HIP;
            set_last_error(st, "canon: nnz readback");
            goto fail;
        }
        if (hipMalloc(&adj, (size_t)(nnz > 0 ? nnz : 1) * sizeof(int32_t)) != hipSuccess) {
            st = TC_ERR_OOM;
            set_last_error(st, "canon: adjacency");
            goto fail;
        }
        if (nnz > 0)
            hipLaunchKernelGGL(k_compact_csr, dim3(grid_for(n, 256, 1 << 16)), dim3(256), 0, s,
                               keys.as<uint64_t>(), flag.as<int32_t>(), pos.as<int64_t>(), n, bits_v,
                               nv, adj, off);
    } else {
        if (hipMalloc(&adj, sizeof(int32_t)) != hipSuccess) {
            st = TC_ERR_OOM;
            goto fail;
        }
    }
    if (hipStreamSynchronize(s) != hipSuccess || hipGetLastError() != hipSuccess) {
        st = TC_ERR_HIP;
        set_last_error(st, "canon: kernels failed");
        goto fail;
    }
    cleanup();
    *off_out = off;
    *adj_out = adj;
    *nnz_out = nnz;
    return TC_OK;
fail:
    cleanup();
    if (off) (void)hipFree(off);
    if (adj) (void)hipFree(adj);
    return st;
}

}  // namespace tc
