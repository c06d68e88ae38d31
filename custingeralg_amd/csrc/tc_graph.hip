// tc_graph.hip -- the dynamic adjacency store: creation from CSR, validation,
// snapshots.  Stands in for the cuStinger container the reference builds with
// initializeCuStinger(eInitStateCSR) (tests/cct.cpp:261-277) and reads through
// dVD->getUsed()/getAdj() (src/static_triangle_counting/cct.cu:235-261).
#include "tc_internal.hpp"

#include <algorithm>
#include <mutex>

namespace tc {

thread_local Error g_last = {TC_OK, ""};

void set_last_error(int code, const std::string &msg)
{
    g_last.code = code;
    g_last.msg = msg;
}

int DevBuf::ensure(size_t need)
{
    if (need == 0) need = 16;
    if (bytes >= need) return TC_OK;
    // a buffer that has to grow again gets 25% headroom: batch-sized scratch
    // varies from call to call, and every regrowth stalls the device (free,
    // then allocate); a first allocation is exact
    if (p) {
        (void)hipFree(p);
        p = nullptr;
        bytes = 0;
        need += need / 4;
    }
    hipError_t e = hipMalloc(&p, need);
    if (e != hipSuccess) {
        p = nullptr;
        (void)hipGetLastError();
        set_last_error(TC_ERR_OOM, "hipMalloc of " + std::to_string(need) + " bytes failed");
        return TC_ERR_OOM;
    }
    bytes = need;
    return TC_OK;
}

void DevBuf::release()
{
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
}

int Graph::ensure_events()
{
    if (events_ready) return TC_OK;
    if (!hcounts)
        TC_HIP(hipHostMalloc(reinterpret_cast<void **>(&hcounts), 32 * sizeof(unsigned long long),
                             hipHostMallocDefault));
    for (auto &e : ev) TC_HIP(hipEventCreate(&e));
    events_ready = true;
    return TC_OK;
}

void Graph::drop_capture()
{
    if (cexec) (void)hipGraphExecDestroy(cexec);
    if (cgraph) (void)hipGraphDestroy(cgraph);
    cexec = nullptr;
    cgraph = nullptr;
    cexec_key = 0;
}

namespace {

__global__ void k_len_from_off(const int64_t *__restrict__ off, int32_t nv, int32_t *__restrict__ len,
                               int64_t *__restrict__ boff, int32_t *__restrict__ cap,
                               int32_t slack_pct)
{
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < nv;
         v += (int64_t)gridDim.x * blockDim.x) {
        int64_t l = off[v + 1] - off[v];
        len[v] = (int32_t)l;
        if (slack_pct <= 0) {
            cap[v] = (int32_t)l;
            boff[v] = off[v];
        } else {
            int64_t want = l + (l * slack_pct + 99) / 100;
            int64_t c = 1;
            while (c < want) c <<= 1;
            cap[v] = (int32_t)c;
        }
    }
}

// int32 offsets (the reference's length_t) widened to int64.
__global__ void k_widen_off(const int32_t *__restrict__ in, int64_t n, int64_t *__restrict__ out)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        out[i] = in[i];
}

// Copy row v of a CSR into its block (blocks laid out by an exclusive scan of cap).
__global__ void k_copy_rows(const int64_t *__restrict__ off, const int32_t *__restrict__ adj,
                            const int64_t *__restrict__ boff, int32_t nv, int32_t *__restrict__ pool)
{
    // one wave per row, rows grid-strided
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t wstride = (int64_t)gridDim.x * (blockDim.x / kWave);
    for (int64_t v = (int64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; v < nv;
         v += wstride) {
        int64_t a = off[v], l = off[v + 1] - a, d = boff[v];
        for (int64_t i = lane; i < l; i += kWave) pool[d + i] = adj[a + i];
    }
}

__device__ inline int32_t row_search_g(const int64_t *__restrict__ coff, int64_t e, int32_t lo,
                                       int32_t hi)
{
    while (lo < hi) {
        int32_t mid = lo + ((hi - lo + 1) >> 1);
        if (coff[mid] <= e) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Validation over compact entries: [0] unsorted/dup, [1] out of range,
// [2] self loops, [3] missing reverse entries, [4] max degree.
__global__ void k_validate(const int64_t *__restrict__ coff, const int64_t *__restrict__ boff,
                           const int32_t *__restrict__ len, const int32_t *__restrict__ pool,
                           int32_t nv, int64_t nnz, int check_sym,
                           unsigned long long *__restrict__ res)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t nchunks = ceil_div(nnz, kWave);
    const int64_t wstride = (int64_t)gridDim.x * (blockDim.x / kWave);
    unsigned long long bad = 0, oor = 0, loops = 0, miss = 0;
    for (int64_t c = (int64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; c < nchunks;
         c += wstride) {
        int64_t e = c * kWave + lane;
        int64_t e0 = c * kWave, elast = min(e0 + kWave, nnz) - 1;
        int32_t rf = 0, rl = 0;
        if (lane == 0) rf = row_search_g(coff, e0, 0, nv - 1);
        if (lane == 1) rl = row_search_g(coff, elast, 0, nv - 1);
        rf = __shfl(rf, 0);
        rl = __shfl(rl, 1);
        if (e > elast) continue;
        int32_t a = row_search_g(coff, e, rf, rl);
        int64_t k = e - coff[a];
        int64_t addr = boff[a] + k;
        int32_t b = pool[addr];
        if (b < 0 || b >= nv) {
            ++oor;
            continue;
        }
        if (k > 0 && pool[addr - 1] >= b) ++bad;
        if (b == a) ++loops;
        if (check_sym) {
            const int32_t *lst = pool + boff[b];
            int32_t lo = 0, hi = len[b];
            while (lo < hi) {
                int32_t mid = (lo + hi) >> 1;
                if (lst[mid] < a) lo = mid + 1; else hi = mid;
            }
            if (!(lo < len[b] && lst[lo] == a)) ++miss;
        }
    }
    for (int d = 32; d > 0; d >>= 1) {
        bad += __shfl_xor(bad, d);
        oor += __shfl_xor(oor, d);
        loops += __shfl_xor(loops, d);
        miss += __shfl_xor(miss, d);
    }
    if (lane == 0) {
        if (bad) atomicAdd(&res[0], bad);
        if (oor) atomicAdd(&res[1], oor);
        if (loops) atomicAdd(&res[2], loops);
        if (miss) atomicAdd(&res[3], miss);
    }
}

__global__ void k_max_len(const int32_t *__restrict__ len, int32_t nv, unsigned long long *res)
{
    unsigned long long m = 0;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < nv;
         v += (int64_t)gridDim.x * blockDim.x)
        m = max(m, (unsigned long long)len[v]);
    for (int d = 32; d > 0; d >>= 1) m = max(m, (unsigned long long)__shfl_xor(m, d));
    if ((threadIdx.x & 63) == 0) atomicMax(res, m);
}

__global__ void k_zero_u64(unsigned long long *p, int n)
{
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0;
}

__global__ void k_gather_rows(const int64_t *__restrict__ coff, const int64_t *__restrict__ boff,
                              const int32_t *__restrict__ pool, int32_t nv, int32_t *__restrict__ adj)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t wstride = (int64_t)gridDim.x * (blockDim.x / kWave);
    for (int64_t v = (int64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; v < nv;
         v += wstride) {
        int64_t a = coff[v], l = coff[v + 1] - a, s = boff[v];
        for (int64_t i = lane; i < l; i += kWave) adj[a + i] = pool[s + i];
    }
}

}  // namespace

int validate_graph(Graph &g, bool check_symmetry, int64_t *bad_rows)
{
    const int32_t nv = g.st.nv;
    const int64_t nnz = g.st.nnz;
    hipStream_t s = g.stream;
    TC_TRY(g.accum.ensure(16 * sizeof(unsigned long long)));
    unsigned long long *res = g.accum.as<unsigned long long>();
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, s, res, 16);
    TC_HIP(hipGetLastError());
    if (nv > 0) {
        TC_TRY(g.coff.ensure((size_t)(nv + 1) * sizeof(int64_t)));
        int64_t *coff = g.coff.as<int64_t>();
        TC_TRY(scan_counts_i32(g.st.len.as<int32_t>(), coff, nv, g.tmp, s));
        if (nnz > 0) {
            int64_t nchunks = ceil_div(nnz, kWave);
            hipLaunchKernelGGL(k_validate, dim3(grid_for(nchunks, 4, 1 << 16)), dim3(256), 0, s, coff,
                               g.st.boff.as<int64_t>(), g.st.len.as<int32_t>(),
                               g.st.pool.as<int32_t>(), nv, nnz, check_symmetry ? 1 : 0, res);
            TC_HIP(hipGetLastError());
        }
        hipLaunchKernelGGL(k_max_len, dim3(grid_for(nv, 256, 4096)), dim3(256), 0, s,
                           g.st.len.as<int32_t>(), nv, res + 4);
        TC_HIP(hipGetLastError());
    }
    unsigned long long h[16];
    TC_HIP(hipMemcpyAsync(h, res, sizeof(h), hipMemcpyDeviceToHost, s));
    TC_HIP(hipStreamSynchronize(s));
    if (bad_rows) *bad_rows = (int64_t)(h[0] + h[1]);
    if (h[0] || h[1]) {
        set_last_error(TC_ERR_INVALID,
                       "graph rows must be sorted strictly ascending with ids in [0,nv): " +
                           std::to_string(h[0]) + " unsorted/duplicate, " + std::to_string(h[1]) +
                           " out-of-range entries");
        return TC_ERR_INVALID;
    }
    g.validated = true;
    g.self_loops = (int64_t)h[2];
    if (check_symmetry) g.symmetric = h[3] == 0 ? 1 : 0;
    g.max_degree = (int32_t)h[4];
    return TC_OK;
}

int graph_create_csr(int32_t nv, int64_t nnz, const int64_t *off, const int32_t *adj, int location,
                     uint32_t flags, int32_t slack_pct, Graph **out)
{
    if (!out || nv < 0 || nnz < 0 || (nv > 0 && !off) || (nnz > 0 && !adj) ||
        (location != TC_HOST && location != TC_DEVICE)) {
        set_last_error(TC_ERR_INVALID, "tc_graph_create_csr: bad arguments");
        return TC_ERR_INVALID;
    }
    const bool off32 = (flags & TC_GRAPH_OFFSETS_I32) != 0;
    std::vector<int64_t> host_wide;
    if (off32 && location == TC_HOST && nv > 0) {
        const int32_t *o32 = reinterpret_cast<const int32_t *>(off);
        host_wide.assign(o32, o32 + (size_t)nv + 1);
        off = host_wide.data();
    }
    if (location == TC_HOST && nv > 0) {
        if (off[0] != 0 || off[nv] != nnz) {
            set_last_error(TC_ERR_INVALID, "tc_graph_create_csr: off[0] must be 0 and off[nv] == nnz");
            return TC_ERR_INVALID;
        }
        for (int32_t v = 0; v < nv; ++v)
            if (off[v + 1] < off[v] || off[v + 1] - off[v] > INT32_MAX) {
                set_last_error(TC_ERR_INVALID, "tc_graph_create_csr: offsets must be monotone");
                return TC_ERR_INVALID;
            }
    }
    Graph *g = new Graph();
    g->st.nv = nv;
    g->st.nnz = nnz;
    auto fail = [&](int code) {
        g->st.boff.release();
        g->st.len.release();
        g->st.cap.release();
        g->st.pool.release();
        delete g;
        return code;
    };
    int st;
    size_t nvb = (size_t)std::max<int32_t>(nv, 1);
    if ((st = g->st.boff.ensure(nvb * sizeof(int64_t))) != TC_OK) return fail(st);
    if ((st = g->st.len.ensure(nvb * sizeof(int32_t))) != TC_OK) return fail(st);
    if ((st = g->st.cap.ensure(nvb * sizeof(int32_t))) != TC_OK) return fail(st);
    // device copy of the offsets (temporarily in coff)
    if ((st = g->coff.ensure((nvb + 1) * sizeof(int64_t))) != TC_OK) return fail(st);
    int64_t *doff = g->coff.as<int64_t>();
    hipStream_t s = nullptr;
    if (nv > 0) {
        if (off32 && location == TC_DEVICE) {
            hipLaunchKernelGGL(k_widen_off, dim3(grid_for((int64_t)nv + 1, 256, 65536)), dim3(256), 0, s,
                               reinterpret_cast<const int32_t *>(off), (int64_t)nv + 1, doff);
            if (hipGetLastError() != hipSuccess) {
                set_last_error(TC_ERR_HIP, "widening of the offsets failed");
                return fail(TC_ERR_HIP);
            }
        } else if (hipMemcpy(doff, off, (size_t)(nv + 1) * sizeof(int64_t),
                             location == TC_HOST ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice) !=
                   hipSuccess) {
            set_last_error(TC_ERR_HIP, "copy of offsets failed");
            return fail(TC_ERR_HIP);
        }
        if (location == TC_DEVICE) {
            int64_t ends[2];
            if (hipMemcpy(&ends[0], doff, sizeof(int64_t), hipMemcpyDeviceToHost) != hipSuccess ||
                hipMemcpy(&ends[1], doff + nv, sizeof(int64_t), hipMemcpyDeviceToHost) != hipSuccess) {
                set_last_error(TC_ERR_HIP, "read-back of the offsets failed");
                return fail(TC_ERR_HIP);
            }
            if (ends[0] != 0 || ends[1] != nnz) {
                set_last_error(TC_ERR_INVALID, "tc_graph_create_csr: off[0] must be 0 and off[nv] == nnz");
                return fail(TC_ERR_INVALID);
            }
        }
        hipLaunchKernelGGL(k_len_from_off, dim3(grid_for(nv, 256, 65536)), dim3(256), 0, s, doff, nv,
                           g->st.len.as<int32_t>(), g->st.boff.as<int64_t>(),
                           g->st.cap.as<int32_t>(), slack_pct);
    }
    int64_t pool_n = nnz;
    if (slack_pct > 0 && nv > 0) {
        // block offsets = exclusive scan of capacities
        DevBuf tmp;
        DevBuf scan;
        if ((st = scan.ensure((nvb + 1) * sizeof(int64_t))) != TC_OK) return fail(st);
        if ((st = scan_counts_i32(g->st.cap.as<int32_t>(), scan.as<int64_t>(), nv, tmp, s)) != TC_OK)
            return fail(st);
        if (hipMemcpy(&pool_n, scan.as<int64_t>() + nv, sizeof(int64_t), hipMemcpyDeviceToHost) !=
                hipSuccess ||
            hipMemcpy(g->st.boff.p, scan.p, (size_t)nv * sizeof(int64_t), hipMemcpyDeviceToDevice) !=
                hipSuccess) {
            set_last_error(TC_ERR_HIP, "copy of the block offsets failed");
            return fail(TC_ERR_HIP);
        }
        tmp.release();
        scan.release();
    }
    // pool gets 25% headroom for block moves before a compaction is needed
    int64_t pool_cap = pool_n + pool_n / 4 + 1024;
    if ((st = g->st.pool.ensure((size_t)pool_cap * sizeof(int32_t))) != TC_OK) return fail(st);
    g->st.pool_used = pool_n;
    if (nnz > 0) {
        if (slack_pct <= 0) {
            if (hipMemcpy(g->st.pool.p, adj, (size_t)nnz * sizeof(int32_t),
                          location == TC_HOST ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice) !=
                hipSuccess) {
                set_last_error(TC_ERR_HIP, "copy of adjacency failed");
                return fail(TC_ERR_HIP);
            }
        } else {
            const int32_t *dadj = adj;
            DevBuf stage;
            if (location == TC_HOST) {
                if ((st = stage.ensure((size_t)nnz * sizeof(int32_t))) != TC_OK) return fail(st);
                if (hipMemcpy(stage.p, adj, (size_t)nnz * sizeof(int32_t), hipMemcpyHostToDevice) !=
                    hipSuccess) {
                    set_last_error(TC_ERR_HIP, "copy of adjacency failed");
                    return fail(TC_ERR_HIP);
                }
                dadj = stage.as<int32_t>();
            }
            hipLaunchKernelGGL(k_copy_rows, dim3(grid_for(nv, 4, 65536)), dim3(256), 0, s, doff, dadj,
                               g->st.boff.as<int64_t>(), nv, g->st.pool.as<int32_t>());
            if (hipDeviceSynchronize() != hipSuccess) {
                set_last_error(TC_ERR_HIP, "row copy failed");
                return fail(TC_ERR_HIP);
            }
            stage.release();
        }
    }
    if (hipDeviceSynchronize() != hipSuccess) {
        set_last_error(TC_ERR_HIP, "graph creation kernels failed");
        return fail(TC_ERR_HIP);
    }
    if (!(flags & TC_GRAPH_NO_VALIDATE)) {
        int64_t bad = 0;
        st = validate_graph(*g, !(flags & TC_GRAPH_ASSUME_SYMMETRIC), &bad);
        if (st != TC_OK) return fail(st);
        if (flags & TC_GRAPH_ASSUME_SYMMETRIC) g->symmetric = 1;
    } else if (flags & TC_GRAPH_ASSUME_SYMMETRIC) {
        g->symmetric = 1;
    }
    *out = g;
    return TC_OK;
}

int graph_to_csr(Graph &g, int64_t *off, int32_t *adj, int location)
{
    const int32_t nv = g.st.nv;
    TC_TRY(g.coff.ensure((size_t)(nv + 1) * sizeof(int64_t)));
    int64_t *coff = g.coff.as<int64_t>();
    TC_TRY(scan_counts_i32(g.st.len.as<int32_t>(), coff, nv, g.tmp, g.stream));
    int64_t *doff = coff;
    int32_t *dadj = adj;
    DevBuf stage;
    if (location == TC_HOST) {
        TC_TRY(stage.ensure((size_t)std::max<int64_t>(g.st.nnz, 1) * sizeof(int32_t)));
        dadj = stage.as<int32_t>();
    }
    if (g.st.nnz > 0)
        hipLaunchKernelGGL(k_gather_rows, dim3(grid_for(nv, 4, 65536)), dim3(256), 0, g.stream, coff,
                           g.st.boff.as<int64_t>(), g.st.pool.as<int32_t>(), nv, dadj);
    TC_HIP(hipGetLastError());
    TC_HIP(hipStreamSynchronize(g.stream));
    hipMemcpyKind k = location == TC_HOST ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
    TC_HIP(hipMemcpy(off, doff, (size_t)(nv + 1) * sizeof(int64_t), k));
    if (location == TC_HOST && g.st.nnz > 0)
        TC_HIP(hipMemcpy(adj, dadj, (size_t)g.st.nnz * sizeof(int32_t), hipMemcpyDeviceToHost));
    stage.release();
    return TC_OK;
}

// A read-only view of a caller's device CSR: per-vertex offsets and lengths
// are built, the adjacency array itself is used in place and never freed.
int graph_borrow_csr(int32_t nv, int64_t nnz, const int64_t *d_off, const int32_t *d_adj,
                     hipStream_t s, Graph **out)
{
    Graph *g = new Graph();
    g->st.nv = nv;
    g->st.nnz = nnz;
    g->stream = s;
    g->borrowed = true;
    const size_t nvb = (size_t)std::max<int32_t>(nv, 1);
    int st;
    if ((st = g->st.boff.ensure(nvb * sizeof(int64_t))) != TC_OK ||
        (st = g->st.len.ensure(nvb * sizeof(int32_t))) != TC_OK ||
        (st = g->st.cap.ensure(nvb * sizeof(int32_t))) != TC_OK) {
        graph_destroy(g);
        return st;
    }
    if (nv > 0)
        hipLaunchKernelGGL(k_len_from_off, dim3(grid_for(nv, 256, 65536)), dim3(256), 0, s, d_off, nv,
                           g->st.len.as<int32_t>(), g->st.boff.as<int64_t>(), g->st.cap.as<int32_t>(),
                           0);
    g->st.pool.p = const_cast<int32_t *>(d_adj);
    g->st.pool.bytes = (size_t)nnz * sizeof(int32_t);
    g->st.pool_used = nnz;
    if (hipStreamSynchronize(s) != hipSuccess || hipGetLastError() != hipSuccess) {
        set_last_error(TC_ERR_HIP, "borrowed graph setup failed");
        graph_destroy(g);
        return TC_ERR_HIP;
    }
    *out = g;
    return TC_OK;
}

namespace {

// compareCUS / comparecus (src/streaming_triangle_counting/cct_upd.cu:442-484):
// one wave per vertex compares the used lengths and then the entries of the
// two stores' rows.  res[0] += rows whose lengths differ, res[1] += rows of
// equal length whose entries differ, res[2] = max(nv - v) over differing
// vertices v (0: none), i.e. the first differing vertex.
__global__ void k_compare_rows(int32_t nv, const int64_t *__restrict__ boff_a,
                               const int32_t *__restrict__ len_a, const int32_t *__restrict__ pool_a,
                               const int64_t *__restrict__ boff_b,
                               const int32_t *__restrict__ len_b, const int32_t *__restrict__ pool_b,
                               unsigned long long *__restrict__ res)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t wstride = (int64_t)gridDim.x * (blockDim.x / kWave);
    unsigned long long nsize = 0, nent = 0;
    for (int64_t v = (int64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; v < nv;
         v += wstride) {
        const int32_t la = len_a[v], lb = len_b[v];
        bool diff = la != lb;
        if (diff) {
            nsize += lane == 0;
        } else {
            const int64_t ba = boff_a[v], bb = boff_b[v];
            bool d = false;
            for (int32_t i = lane; i < la; i += kWave) d |= pool_a[ba + i] != pool_b[bb + i];
            diff = __ballot(d) != 0;
            nent += (diff && lane == 0);
        }
        if (diff && lane == 0) atomicMax(&res[2], (unsigned long long)(nv - v));
    }
    const unsigned long long part[2] = {nsize, nent};
    block_add_u64<2>(part, res);
}

}  // namespace

int graph_compare(Graph &a, Graph &b, int64_t *size_mismatch, int64_t *entry_mismatch,
                  int32_t *first_vertex)
{
    if (a.st.nv != b.st.nv) {
        set_last_error(TC_ERR_INVALID, "tc_graph_compare: the stores have different vertex counts");
        return TC_ERR_INVALID;
    }
    hipStream_t s = a.stream;
    TC_HIP(hipStreamSynchronize(b.stream));
    TC_TRY(a.accum.ensure(16 * sizeof(unsigned long long)));
    unsigned long long *res = a.accum.as<unsigned long long>();
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, s, res, 3);
    if (a.st.nv > 0)
        hipLaunchKernelGGL(k_compare_rows, dim3(grid_for(a.st.nv, 4, 65536)), dim3(256), 0, s,
                           a.st.nv, a.st.boff.as<int64_t>(), a.st.len.as<int32_t>(),
                           a.st.pool.as<int32_t>(), b.st.boff.as<int64_t>(), b.st.len.as<int32_t>(),
                           b.st.pool.as<int32_t>(), res);
    TC_HIP(hipGetLastError());
    unsigned long long h[3];
    TC_HIP(hipMemcpyAsync(h, res, sizeof(h), hipMemcpyDeviceToHost, s));
    TC_HIP(hipStreamSynchronize(s));
    *size_mismatch = (int64_t)h[0];
    *entry_mismatch = (int64_t)h[1];
    *first_vertex = h[2] == 0 ? -1 : (int32_t)((int64_t)a.st.nv - (int64_t)h[2]);
    return TC_OK;
}

void graph_destroy(Graph *g)
{
    if (!g) return;
    (void)hipDeviceSynchronize();
    if (g->borrowed) {
        g->st.pool.p = nullptr;
        g->st.pool.bytes = 0;
    }
    DevBuf *bufs[] = {&g->st.boff, &g->st.len, &g->st.cap, &g->st.pool, &g->vinfo, &g->padj,
                      &g->coff, &g->crow, &g->work, &g->jobs, &g->tmp, &g->tmp2, &g->tmp3, &g->tmp4,
                      &g->accum, &g->bsrc, &g->bdst, &g->bflag, &g->bkeys, &g->bkeys2,
                      &g->bvals2, &g->boffb, &g->bdel, &g->flagpool, &g->vmark, &g->order,
                      &g->rnk, &g->tasks, &g->pcnt, &g->hjobs, &g->outr};
    for (DevBuf *b : bufs) b->release();
    for (DevBuf &b : g->sb) b.release();
    for (DevBuf &b : g->ov) b.release();
    for (DevBuf &b : g->lbuf) b.release();
    if (g->events_ready)
        for (auto &e : g->ev) (void)hipEventDestroy(e);
    if (g->hcounts) (void)hipHostFree(g->hcounts);
    g->drop_capture();
    if (g->bexec) (void)hipGraphExecDestroy(g->bexec);
    if (g->bgraph) (void)hipGraphDestroy(g->bgraph);
    if (g->cap_stream) (void)hipStreamDestroy(g->cap_stream);
    delete g;
}

}  // namespace tc
