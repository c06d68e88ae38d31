// tc_prims.hip -- device-wide exclusive scan and LSD radix sort, written for
// 64-lane wavefronts.  These replace the moderngpu scan / sorted_search and
// cub BlockRadixSort the reference links against (include/load_balance.cuh:81,100;
// include/operators.cuh:175,203).
#include "tc_internal.hpp"

namespace tc {

// ------------------------------------------------------------------ scan --
namespace {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;

__device__ inline int64_t wave_incl_scan(int64_t x)
{
    const int lane = threadIdx.x & (kWave - 1);
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        int64_t y = __shfl_up(x, d, kWave);
        if (lane >= d) x += y;
    }
    return x;
}

// Block-wide exclusive scan of one value per thread; returns the block total.
__device__ inline int64_t block_excl_scan(int64_t x, int64_t &excl, int64_t *lds_wave)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = threadIdx.x / kWave;
    const int nw = blockDim.x / kWave;
    int64_t incl = wave_incl_scan(x);
    if (lane == kWave - 1) lds_wave[wid] = incl;
    __syncthreads();
    int64_t before = 0, total = 0;
    for (int w = 0; w < nw; ++w) {
        int64_t s = lds_wave[w];
        if (w < wid) before += s;
        total += s;
    }
    excl = before + incl - x;
    __syncthreads();
    return total;
}

template <class T>
__global__ void k_tile_reduce(const T *__restrict__ in, int64_t n, int64_t *__restrict__ sums)
{
    __shared__ int64_t lw[kScanThreads / kWave];
    int64_t base = (int64_t)blockIdx.x * kScanTile;
    int64_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        int64_t i = base + (int64_t)k * kScanThreads + threadIdx.x;
        if (i < n) s += (int64_t)in[i];
    }
    int64_t excl;
    int64_t tot = block_excl_scan(s, excl, lw);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

// Scans one tile; thread t owns the contiguous items [t*8, t*8+8) of the tile.
template <class T>
__global__ void k_tile_scan(const T *__restrict__ in, int64_t *__restrict__ out, int64_t n,
                            const int64_t *__restrict__ tile_off, int64_t *__restrict__ total)
{
    __shared__ int64_t lw[kScanThreads / kWave];
    __shared__ int64_t stage[kScanTile];
    int64_t base = (int64_t)blockIdx.x * kScanTile;
    // coalesced load into LDS
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        int j = k * kScanThreads + threadIdx.x;
        int64_t i = base + j;
        stage[j] = (i < n) ? (int64_t)in[i] : 0;
    }
    __syncthreads();
    int64_t v[kScanItems];
    int64_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        v[k] = stage[threadIdx.x * kScanItems + k];
        s += v[k];
    }
    int64_t excl;
    int64_t tot = block_excl_scan(s, excl, lw);
    int64_t off = (tile_off ? tile_off[blockIdx.x] : 0) + excl;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        stage[threadIdx.x * kScanItems + k] = off;
        off += v[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        int j = k * kScanThreads + threadIdx.x;
        int64_t i = base + j;
        if (i < n) out[i] = stage[j];
    }
    if (total && blockIdx.x == gridDim.x - 1 && threadIdx.x == 0)
        *total = (tile_off ? tile_off[blockIdx.x] : 0) + tot;
}

__global__ void k_zero_total(int64_t *total) { *total = 0; }

// Scans of at most kScanSmall items: one block, thread t holds items
// t + 1024k (every load and store is a full coalesced row), four wave scans
// and one barrier.  Small update batches scan a dozen such arrays, where a
// launch's latency is all it costs.
constexpr int kScanSmallThreads = 1024, kScanSmallRows = 4;
constexpr int kScanSmall = kScanSmallThreads * kScanSmallRows;

template <class T>
__global__ __launch_bounds__(kScanSmallThreads) void k_scan_small(const T *__restrict__ in,
                                                                  int64_t *__restrict__ out, int64_t n,
                                                                  int64_t *__restrict__ total)
{
    __shared__ int64_t ws[kScanSmallRows][kScanSmallThreads / kWave];
    const int t = threadIdx.x, lane = t & (kWave - 1), wid = t / kWave;
    int64_t v[kScanSmallRows], incl[kScanSmallRows];
#pragma unroll
    for (int k = 0; k < kScanSmallRows; ++k) {
        const int64_t i = (int64_t)k * kScanSmallThreads + t;
        v[k] = i < n ? (int64_t)in[i] : 0;
    }
#pragma unroll
    for (int k = 0; k < kScanSmallRows; ++k) {
        incl[k] = wave_incl_scan(v[k]);
        if (lane == kWave - 1) ws[k][wid] = incl[k];
    }
    __syncthreads();
    int64_t base = 0;
#pragma unroll
    for (int k = 0; k < kScanSmallRows; ++k) {
        int64_t before = 0, row = 0;
        for (int w = 0; w < kScanSmallThreads / kWave; ++w) {
            const int64_t x = ws[k][w];
            before += w < wid ? x : 0;
            row += x;
        }
        const int64_t i = (int64_t)k * kScanSmallThreads + t;
        if (i < n) out[i] = base + before + incl[k] - v[k];
        base += row;
    }
    if (total && t == 0) *total = base;
}

template <class T>
int scan_impl(const T *in, int64_t *out, int64_t n, int64_t *total, DevBuf &tmp, hipStream_t s,
              size_t tmp_base)
{
    if (n <= 0) {
        // a kernel, not a memset: counts may be recorded into a hipGraph
        if (total) hipLaunchKernelGGL(k_zero_total, dim3(1), dim3(1), 0, s, total);
        TC_HIP(hipGetLastError());
        return TC_OK;
    }
    if (n <= kScanSmall) {  // (in place too: a thread writes only what it read)
        hipLaunchKernelGGL(k_scan_small<T>, dim3(1), dim3(kScanSmallThreads), 0, s, in, out, n, total);
        TC_HIP(hipGetLastError());
        return TC_OK;
    }
    int64_t tiles = ceil_div(n, kScanTile);
    if (tiles == 1) {
        hipLaunchKernelGGL(k_tile_scan<T>, dim3(1), dim3(kScanThreads), 0, s, in, out, n,
                           (const int64_t *)nullptr, total);
        TC_HIP(hipGetLastError());
        return TC_OK;
    }
    // scratch layout: [tmp_base, tmp_base + tiles) holds tile sums
    size_t need = (tmp_base + (size_t)tiles + 1) * sizeof(int64_t);
    if (tmp.bytes < need) {
        set_last_error(TC_ERR_INVALID, "scan scratch too small");
        return TC_ERR_INVALID;
    }
    int64_t *sums = tmp.as<int64_t>() + tmp_base;
    hipLaunchKernelGGL(k_tile_reduce<T>, dim3((unsigned)tiles), dim3(kScanThreads), 0, s, in, n,
                       sums);
    TC_HIP(hipGetLastError());
    TC_TRY(scan_impl<int64_t>(sums, sums, tiles, nullptr, tmp, s, tmp_base + (size_t)tiles + 1));
    hipLaunchKernelGGL(k_tile_scan<T>, dim3((unsigned)tiles), dim3(kScanThreads), 0, s, in, out,
                       n, sums, total);
    TC_HIP(hipGetLastError());
    return TC_OK;
}

size_t scan_scratch_words(int64_t n)
{
    size_t w = 0;
    while (n > kScanTile) {
        n = ceil_div(n, kScanTile);
        w += (size_t)n + 1;
    }
    return w + 1;
}

}  // namespace

int scan_exclusive_i64(const int64_t *in, int64_t *out, int64_t n, int64_t *total, DevBuf &tmp,
                       hipStream_t s)
{
    TC_TRY(tmp.ensure(scan_scratch_words(n) * sizeof(int64_t)));
    return scan_impl<int64_t>(in, out, n, total, tmp, s, 0);
}

int scan_counts_i32(const int32_t *in, int64_t *out, int64_t n, DevBuf &tmp, hipStream_t s)
{
    TC_TRY(tmp.ensure(scan_scratch_words(n) * sizeof(int64_t)));
    return scan_impl<int32_t>(in, out, n, out + n, tmp, s, 0);
}

// ------------------------------------------------------------ radix sort --
namespace {

constexpr int kSortThreads = 256;
constexpr int kSortRounds = 16;
constexpr int kSortTile = kSortThreads * kSortRounds;  // 4096 keys per block
constexpr int kRadix = 256;

__global__ void k_radix_hist(const uint64_t *__restrict__ keys, int64_t n, int shift,
                             int64_t *__restrict__ hist, int64_t nblocks)
{
    __shared__ uint32_t h[kRadix];
    h[threadIdx.x] = 0;
    __syncthreads();
    int64_t base = (int64_t)blockIdx.x * kSortTile;
    for (int r = 0; r < kSortRounds; ++r) {
        int64_t i = base + (int64_t)r * kSortThreads + threadIdx.x;
        if (i < n) atomicAdd(&h[(uint32_t)(keys[i] >> shift) & (kRadix - 1)], 1u);
    }
    __syncthreads();
    // digit-major: all blocks of digit 0, then digit 1, ...
    hist[(int64_t)threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

// Stable scatter: keys of a tile are ranked in tile order, round by round.
__global__ void k_radix_scatter(const uint64_t *__restrict__ keys_in,
                                const int32_t *__restrict__ vals_in, int64_t n, int shift,
                                const int64_t *__restrict__ hist_scan, int64_t nblocks,
                                uint64_t *__restrict__ keys_out, int32_t *__restrict__ vals_out)
{
    constexpr int NW = kSortThreads / kWave;
    __shared__ int64_t base_off[kRadix];
    __shared__ uint32_t running[kRadix];
    __shared__ uint32_t wcnt[NW][kRadix];
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = threadIdx.x / kWave;
    base_off[threadIdx.x] = hist_scan[(int64_t)threadIdx.x * nblocks + blockIdx.x];
    running[threadIdx.x] = 0;
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (kWave - lane));
    int64_t base = (int64_t)blockIdx.x * kSortTile;
    for (int r = 0; r < kSortRounds; ++r) {
#pragma unroll
        for (int w = 0; w < NW; ++w) wcnt[w][threadIdx.x] = 0;
        __syncthreads();
        int64_t i = base + (int64_t)r * kSortThreads + threadIdx.x;
        bool valid = i < n;
        uint64_t k = valid ? keys_in[i] : 0;
        int32_t pv = (valid && vals_in) ? vals_in[i] : 0;
        uint32_t d = (uint32_t)(k >> shift) & (kRadix - 1);
        // peers: lanes of this wave holding the same digit (valid only)
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            uint64_t bal = __ballot(valid && ((d >> b) & 1u));
            peers &= ((d >> b) & 1u) ? bal : ~bal;
        }
        uint32_t rank_w = (uint32_t)__popcll(peers & lt_mask);
        uint32_t cnt_w = (uint32_t)__popcll(peers);
        if (valid && rank_w == 0) wcnt[wid][d] = cnt_w;
        __syncthreads();
        if (valid) {
            uint32_t before = 0;
            for (int w = 0; w < wid; ++w) before += wcnt[w][d];
            int64_t pos = base_off[d] + running[d] + before + rank_w;
            keys_out[pos] = k;
            if (vals_out) vals_out[pos] = pv;
        }
        __syncthreads();
        uint32_t tot = 0;
#pragma unroll
        for (int w = 0; w < NW; ++w) tot += wcnt[w][threadIdx.x];
        running[threadIdx.x] += tot;
        __syncthreads();
    }
}

}  // namespace

namespace {

// Keys-only sorts of at most kSmallSort keys: one block, a bitonic network in
// LDS over the next power of two (padded with ~0, past every key).  One
// launch instead of three per 8-bit radix pass: small update batches sort
// twice, and their launches, not their keys, are what they cost.
constexpr int kSmallSort = 8192;

__global__ __launch_bounds__(1024) void k_sort_small(uint64_t *__restrict__ keys, int64_t n)
{
    __shared__ uint64_t sk[kSmallSort];
    int m = 1;
    while (m < n) m <<= 1;
    for (int i = threadIdx.x; i < m; i += blockDim.x) sk[i] = i < n ? keys[i] : ~0ull;
    __syncthreads();
    for (int k = 2; k <= m; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < m; i += blockDim.x) {
                const int l = i ^ j;
                if (l > i) {
                    const uint64_t a = sk[i], b = sk[l];
                    const bool up = (i & k) == 0;
                    if ((a > b) == up) {
                        sk[i] = b;
                        sk[l] = a;
                    }
                }
            }
            __syncthreads();
        }
    }
    for (int i = threadIdx.x; i < n; i += blockDim.x) keys[i] = sk[i];
}

// Keys-only sorts of at most kRegSort keys: the same bitonic network with the
// keys in registers (block_bitonic_2048).
__global__ __launch_bounds__(1024) void k_sort_reg(uint64_t *__restrict__ keys, int64_t n)
{
    __shared__ uint64_t buf[2][kRegSort];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int i0 = 128 * w + lane, i1 = i0 + 64;
    uint64_t v0 = i0 < n ? keys[i0] : ~0ull, v1 = i1 < n ? keys[i1] : ~0ull;
    block_bitonic_2048(v0, v1, buf);
    if (i0 < n) keys[i0] = v0;
    if (i1 < n) keys[i1] = v1;
}

}  // namespace

int radix_sort_u64(uint64_t *keys, int32_t *vals, int64_t n, int bits, DevBuf &alt_keys,
                   DevBuf &alt_vals, DevBuf &tmp, hipStream_t s)
{
    if (n <= 1 || bits <= 0) return TC_OK;
    if (!vals && n <= kRegSort) {  // (keys below 2^bits: order by the whole key)
        hipLaunchKernelGGL(k_sort_reg, dim3(1), dim3(1024), 0, s, keys, n);
        TC_HIP(hipGetLastError());
        return TC_OK;
    }
    if (!vals && n <= kSmallSort) {
        hipLaunchKernelGGL(k_sort_small, dim3(1), dim3(1024), 0, s, keys, n);
        TC_HIP(hipGetLastError());
        return TC_OK;
    }
    int64_t nblocks = ceil_div(n, kSortTile);
    TC_TRY(alt_keys.ensure((size_t)n * sizeof(uint64_t)));
    if (vals) TC_TRY(alt_vals.ensure((size_t)n * sizeof(int32_t)));
    int64_t hn = nblocks * kRadix;
    size_t scan_words = scan_scratch_words(hn);
    TC_TRY(tmp.ensure((size_t)(hn + scan_words) * sizeof(int64_t)));
    int64_t *hist = tmp.as<int64_t>();
    // the scan's own scratch lives after the histogram
    DevBuf scan_tmp;
    scan_tmp.p = tmp.as<int64_t>() + hn;
    scan_tmp.bytes = scan_words * sizeof(int64_t);

    uint64_t *ka = keys, *kb = alt_keys.as<uint64_t>();
    int32_t *va = vals, *vb = vals ? alt_vals.as<int32_t>() : nullptr;
    int passes = (bits + 7) / 8;
    for (int p = 0; p < passes; ++p) {
        int shift = p * 8;
        hipLaunchKernelGGL(k_radix_hist, dim3((unsigned)nblocks), dim3(kSortThreads), 0, s, ka,
                           n, shift, hist, nblocks);
        TC_HIP(hipGetLastError());
        TC_TRY(scan_impl<int64_t>(hist, hist, hn, nullptr, scan_tmp, s, 0));
        hipLaunchKernelGGL(k_radix_scatter, dim3((unsigned)nblocks), dim3(kSortThreads), 0, s,
                           ka, va, n, shift, hist, nblocks, kb, vb);
        TC_HIP(hipGetLastError());
        std::swap(ka, kb);
        std::swap(va, vb);
    }
    if (ka != keys) {
        TC_HIP(hipMemcpyAsync(keys, ka, (size_t)n * sizeof(uint64_t), hipMemcpyDeviceToDevice, s));
        if (vals)
            TC_HIP(hipMemcpyAsync(vals, va, (size_t)n * sizeof(int32_t), hipMemcpyDeviceToDevice,
                                  s));
    }
    return TC_OK;
}

}  // namespace tc
