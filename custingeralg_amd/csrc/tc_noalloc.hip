// tc_noalloc.hip -- the CSR-convention count (callDeviceAllTrianglesCSR's
// output) for callers that cannot allocate device memory: code under HIP
// graph capture, tight serving loops, harnesses that forbid hipMalloc.
//
// tc_csr_count_device builds a store and a job plan over the caller's CSR,
// all of it allocated per call.  This variant reads the caller's CSR in place
// and keeps its few counters in `scratch`, a file-scope device array that
// every GPU holds once for the life of the process:
//
//   word 0  error flags of the validation pass (bit 0: offsets not
//           non-decreasing or outside [0, nnz]; bit 1: an id outside
//           [0, nv); bit 2: a row not strictly increasing)
//   word 1  cursor of the 64-entry chunks the count kernel's waves take
//   word 2  the global sum
//
// Its address on each device comes from hipGetSymbolAddress at the first
// call on that device; host code never passes the array's name itself.
// Calls that overlap on one GPU share these words and race.
//
// Count (tests/cct_csr.cu:228-276 for the output; src/static_triangle_counting/
// cct.cu:247-255 for the filters): out[src] = sum over entries (src, dst) of
// the row with dst > src, deg(src) >= 2, deg(dst) >= 2 of |N(src) n N(dst)|,
// modulo 2^32; the global is the exact uint64 sum (3T on a symmetric simple
// graph).  Each wave takes 64 consecutive entries of ind at a time from the
// cursor and intersects each kept entry's two rows with the whole wave: a
// merge path over 64-entry windows of both rows when their lengths are within
// a factor of 16, else a binary search of the shorter row's entries in the
// longer one, one entry per lane.
#include "tc_internal.hpp"

#include <mutex>

namespace tc {
namespace {

constexpr int kNaWords = 32;          // 256 bytes of scratch per device
constexpr int kNaBlock = 256;         // 4 waves per block
constexpr int kNaMaxDevices = 64;

__device__ __align__(256) unsigned long long scratch[kNaWords];

std::mutex g_na_mu;
unsigned long long *g_na_ptr[kNaMaxDevices];  // scratch's address on each device

// The device address of `scratch` on the current device (looked up once per
// device, with that device current).
int scratch_ptr(unsigned long long **out)
{
    int dev = 0;
    TC_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kNaMaxDevices) {
        set_last_error(TC_ERR_UNSUPPORTED, "tc_csr_count_device_noalloc: device index out of range");
        return TC_ERR_UNSUPPORTED;
    }
    std::lock_guard<std::mutex> lk(g_na_mu);
    if (!g_na_ptr[dev]) {
        void *p = nullptr;
        TC_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_na_ptr[dev] = static_cast<unsigned long long *>(p);
    }
    *out = g_na_ptr[dev];
    return TC_OK;
}

__global__ void k_na_clear(unsigned long long *scr, uint32_t *out, int32_t nv)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < kNaWords) scr[i] = 0ull;
    for (int64_t v = i; v < nv; v += (int64_t)gridDim.x * blockDim.x) out[v] = 0u;
}

// Offsets non-decreasing within [0, nnz], ids in range, rows strictly
// increasing (one wave per row for the last check).
__global__ void k_na_validate(int32_t nv, int64_t nnz, const int64_t *__restrict__ off,
                              const int32_t *__restrict__ ind, unsigned long long *scr)
{
    unsigned flags = 0;
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t i0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (int64_t v = i0; v < nv; v += stride) {
        const int64_t a = off[v], b = off[v + 1];
        if (a < 0 || b < a || b > nnz) flags |= 1u;
    }
    if (i0 == 0 && (off[0] != 0 || off[nv] != nnz)) flags |= 1u;
    for (int64_t e = i0; e < nnz; e += stride) {
        const int32_t x = ind[e];
        if (x < 0 || x >= nv) flags |= 2u;
    }
    for (int64_t v = i0 / kWave; v < nv; v += stride / kWave) {
        const int64_t a = off[v], b = off[v + 1];
        if (a < 0 || b < a || b > nnz) continue;
        for (int64_t k = a + 1 + lane; k < b; k += kWave)
            if (ind[k] <= ind[k - 1]) flags |= 4u;
    }
    if (flags) atomicOr(&scr[0], (unsigned long long)flags);
}

__device__ inline int64_t readlane64(int64_t x, int lane)
{
    const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)(uint64_t)x, lane);
    const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)((uint64_t)x >> 32), lane);
    return (int64_t)(((uint64_t)hi << 32) | lo);
}

// |X n Y| of two sorted rows by the whole wave (wave-uniform arguments).
// Merge path over 64-entry windows: values up to m = min(last x, last y) are
// settled by the two windows, so a pair x == y <= m is counted there and
// nowhere else; each lane finds its x's co-rank in the Y window by a
// six-step search over the window's lanes.  The window whose last entry is m
// advances by its whole length, the other past its entries <= m, so every
// step makes progress even on rows that are not sorted.
__device__ inline uint32_t na_merge_path(const int32_t *__restrict__ X, int64_t nx,
                                         const int32_t *__restrict__ Y, int64_t ny)
{
    const int lane = threadIdx.x & (kWave - 1);
    int64_t ix = 0, iy = 0;
    uint32_t h = 0;
    while (ix < nx && iy < ny) {
        const int32_t cx = (int32_t)min<int64_t>(kWave, nx - ix);
        const int32_t cy = (int32_t)min<int64_t>(kWave, ny - iy);
        const int32_t x = lane < cx ? X[ix + lane] : INT32_MAX;
        const int32_t y = lane < cy ? Y[iy + lane] : INT32_MAX;
        const int32_t lx = __builtin_amdgcn_readlane(x, cx - 1);
        const int32_t ly = __builtin_amdgcn_readlane(y, cy - 1);
        const int32_t m = min(lx, ly);
        int32_t r = 0;
#pragma unroll
        for (int st = kWave / 2; st > 0; st >>= 1) {
            const int32_t p = __shfl(y, r + st - 1, kWave);
            r += (p < x) ? st : 0;
        }
        const int32_t yr = __shfl(y, r, kWave);  // every lane: ds_bpermute reads active lanes only
        const bool match = lane < cx && x <= m && yr == x;
        h += (uint32_t)__popcll(__ballot(match));
        ix += (lx == m) ? cx : __popcll(__ballot(lane < cx && x <= m));
        iy += (ly == m) ? cy : __popcll(__ballot(lane < cy && y <= m));
    }
    return h;
}

// |S n Lg| with S much shorter than Lg: each lane searches one entry of S in
// Lg at a time (lower bound), the wave covering S in 64-entry strides.
__device__ inline uint32_t na_search(const int32_t *__restrict__ S, int64_t ns,
                                     const int32_t *__restrict__ Lg, int64_t nl)
{
    const int lane = threadIdx.x & (kWave - 1);
    uint32_t h = 0;
    for (int64_t k0 = 0; k0 < ns; k0 += kWave) {
        const int64_t k = k0 + lane;
        bool hit = false;
        if (k < ns) {
            const int32_t x = S[k];
            int64_t lo = 0, n = nl;
            while (n > 0) {
                const int64_t half = n >> 1;
                if (Lg[lo + half] < x) {
                    lo += half + 1;
                    n -= half + 1;
                } else {
                    n = half;
                }
            }
            hit = lo < nl && Lg[lo] == x;
        }
        h += (uint32_t)__popcll(__ballot(hit));
    }
    return h;
}

__global__ __launch_bounds__(kNaBlock) void k_na_count(int32_t nv, int64_t nnz,
                                                        const int64_t *__restrict__ off,
                                                        const int32_t *__restrict__ ind,
                                                        uint32_t *__restrict__ out,
                                                        unsigned long long *scr)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t nchunks = (nnz + kWave - 1) / kWave;
    unsigned long long total = 0;  // wave-uniform
    for (;;) {
        int64_t c = 0;
        if (lane == 0) c = (int64_t)atomicAdd(&scr[1], 1ull);
        c = readlane64(c, 0);
        if (c >= nchunks) break;
        const int64_t e = c * kWave + lane;
        bool keep = false;
        int32_t src = 0, dst = 0;
        if (e < nnz) {
            // src = the last row starting at or before e (empty rows share offsets)
            int32_t lo = 0, hi = nv;  // off[lo] <= e < off[hi]
            while (hi - lo > 1) {
                const int32_t mid = lo + ((hi - lo) >> 1);
                if (off[mid] <= e) lo = mid; else hi = mid;
            }
            src = lo;
            dst = ind[e];
            keep = dst > src && off[src + 1] - off[src] >= 2 && off[dst + 1] - off[dst] >= 2;
        }
        for (uint64_t kb = __ballot(keep); kb; kb &= kb - 1) {
            const int j = __ffsll((unsigned long long)kb) - 1;
            const int32_t s = __builtin_amdgcn_readlane(src, j);
            const int32_t d = __builtin_amdgcn_readlane(dst, j);
            const int64_t sb = off[s], sl = off[s + 1] - sb;
            const int64_t db = off[d], dl = off[d + 1] - db;
            const int64_t sh = min(sl, dl), lg = max(sl, dl);
            const int32_t *Sh = ind + (sl <= dl ? sb : db);
            const int32_t *Lg = ind + (sl <= dl ? db : sb);
            const uint32_t h = lg > 16 * sh ? na_search(Sh, sh, Lg, lg) : na_merge_path(Sh, sh, Lg, lg);
            if (h) {
                if (lane == 0) atomicAdd(&out[s], h);
                total += h;
            }
        }
    }
    if (lane == 0 && total) atomicAdd(&scr[2], total);
}

}  // namespace

int csr_count_noalloc(int32_t nv, int64_t nnz, const int64_t *off, const int32_t *ind,
                      uint32_t *out, uint64_t *global, hipStream_t s)
{
    unsigned long long *scr = nullptr;
    TC_TRY(scratch_ptr(&scr));
    int dev = 0, cus = 0;
    TC_HIP(hipGetDevice(&dev));
    TC_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const int gclear = grid_for(std::max<int64_t>(nv, kNaWords), kNaBlock, 4096);
    hipLaunchKernelGGL(k_na_clear, dim3(gclear), dim3(kNaBlock), 0, s, scr, out, nv);
    TC_HIP(hipGetLastError());
    if (nv == 0) {
        TC_HIP(hipStreamSynchronize(s));
        if (global) *global = 0;
        return TC_OK;
    }
    const int gval = grid_for(std::max<int64_t>(nv, nnz), kNaBlock, 8192);
    hipLaunchKernelGGL(k_na_validate, dim3(gval), dim3(kNaBlock), 0, s, nv, nnz, off, ind, scr);
    TC_HIP(hipGetLastError());
    unsigned long long flags = 0;
    TC_HIP(hipMemcpyAsync(&flags, scr, sizeof(flags), hipMemcpyDeviceToHost, s));
    TC_HIP(hipStreamSynchronize(s));
    if (flags) {
        set_last_error(TC_ERR_INVALID,
                       std::string("tc_csr_count_device_noalloc: ") +
                           ((flags & 1u) ? "offsets not non-decreasing in [0, nnz]"
                            : (flags & 2u) ? "vertex id out of range"
                                           : "row not strictly increasing"));
        return TC_ERR_INVALID;
    }
    // persistent grid: eight 4-wave blocks per CU
    const int64_t nchunks = (nnz + kWave - 1) / kWave;
    const int gcount = (int)std::max<int64_t>(1, std::min<int64_t>((int64_t)cus * 8,
                                                                    ceil_div(nchunks, kNaBlock / kWave)));
    hipLaunchKernelGGL(k_na_count, dim3(gcount), dim3(kNaBlock), 0, s, nv, nnz, off, ind, out, scr);
    TC_HIP(hipGetLastError());
    unsigned long long g = 0;
    TC_HIP(hipMemcpyAsync(&g, scr + 2, sizeof(g), hipMemcpyDeviceToHost, s));
    TC_HIP(hipStreamSynchronize(s));
    if (global) *global = g;
    return TC_OK;
}

}  // namespace tc

// The C-ABI entry (declared in include/custinger_tc.h).
extern "C" int tc_csr_count_device_noalloc(int32_t nv, int64_t nnz, const int64_t *d_off,
                                           const int32_t *d_ind, uint32_t *out_dev,
                                           uint64_t *global_host, void *hip_stream)
{
    try {
        if (nv < 0 || nnz < 0 || (nv > 0 && !d_off) || (nnz > 0 && !d_ind) ||
            (nv > 0 && !out_dev)) {
            tc::set_last_error(TC_ERR_INVALID, "tc_csr_count_device_noalloc: bad arguments");
            return TC_ERR_INVALID;
        }
        if (nv == 0 && nnz > 0) {
            tc::set_last_error(TC_ERR_INVALID, "tc_csr_count_device_noalloc: entries without vertices");
            return TC_ERR_INVALID;
        }
        return tc::csr_count_noalloc(nv, nnz, d_off, d_ind, out_dev, global_host,
                                     (hipStream_t)hip_stream);
    } catch (...) {
        tc::set_last_error(TC_ERR_INVALID, "tc_csr_count_device_noalloc: unexpected exception");
        return TC_ERR_INVALID;
    }
}
