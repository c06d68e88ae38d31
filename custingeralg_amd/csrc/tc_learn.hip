// tc_learn.hip -- the learned strategy selector and the schedule autotuner.
//
// * Timing collection: sampled oriented tasks (owner b, task a) are run with
//   each intersection strategy in isolation, one wavefront per sample, timed
//   with the shader clock (s_memtime); the three hit counts are cross-checked.
// * Cost model: a 4 -> 32 -> 32 -> 3 ReLU MLP predicting log2(cycles) of the
//   three strategies from (log2|owner|/16, log2|task|/16, log2 ratio/16, bin/2).
//   Training is FP32 on the device: a forward+backward kernel writes per-block
//   gradient partials (no atomics), a second kernel sums them in block order
//   and applies Adam, so a step is deterministic.
// * Inference: the weights are rounded to BF16 (round to nearest even) and a
//   kernel evaluates the network over the 32x32 grid of log2 buckets of
//   (|owner tile|, |task list|), accumulating in FP32, and stores the arg-min
//   strategy per cell.  The count kernel reads that table (TC_POLICY_LEARNED);
//   every strategy gives the same counts, so the table only moves time.
// * Autotuner: the reference sweeps (blocks x blockdim x tpi) and keeps the
//   fastest (tests/cct.cpp:177-209); here the sweep is over bin thresholds and
//   job grains of the static count, and the winner becomes the default.
#include "tc_internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>

namespace tc {

namespace {

constexpr int IN = TC_MLP_IN, HID = TC_MLP_HIDDEN, OUT = TC_MLP_OUT, NP = TC_MLP_PARAMS;
constexpr int oW1 = 0, oB1 = oW1 + HID * IN, oW2 = oB1 + HID, oB2 = oW2 + HID * HID,
              oW3 = oB2 + HID, oB3 = oW3 + OUT * HID;
static_assert(oB3 + OUT == NP, "parameter layout");
constexpr int kTableDim = 32;
constexpr int kTrainBlock = 64;  // samples per block of the backward kernel

// ------------------------------------------------------------------ MLP --

__device__ inline void mlp_forward_f32(const float *P, const float x[IN], float h1[HID],
                                       float h2[HID], float y[OUT])
{
#pragma unroll 4
    for (int r = 0; r < HID; ++r) {
        float a = P[oB1 + r];
#pragma unroll
        for (int c = 0; c < IN; ++c) a = fmaf(P[oW1 + r * IN + c], x[c], a);
        h1[r] = a > 0.f ? a : 0.f;
    }
    for (int r = 0; r < HID; ++r) {
        float a = P[oB2 + r];
#pragma unroll 8
        for (int c = 0; c < HID; ++c) a = fmaf(P[oW2 + r * HID + c], h1[c], a);
        h2[r] = a > 0.f ? a : 0.f;
    }
#pragma unroll
    for (int r = 0; r < OUT; ++r) {
        float a = P[oB3 + r];
        for (int c = 0; c < HID; ++c) a = fmaf(P[oW3 + r * HID + c], h2[c], a);
        y[r] = a;
    }
}

__global__ void k_mlp_forward(const float *__restrict__ params, const float *__restrict__ feat,
                              int64_t n, float *__restrict__ pred)
{
    __shared__ float P[NP];
    for (int i = threadIdx.x; i < NP; i += blockDim.x) P[i] = params[i];
    __syncthreads();
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < n;
         s += (int64_t)gridDim.x * blockDim.x) {
        float x[IN], h1[HID], h2[HID], y[OUT];
#pragma unroll
        for (int c = 0; c < IN; ++c) x[c] = feat[s * IN + c];
        mlp_forward_f32(P, x, h1, h2, y);
#pragma unroll
        for (int r = 0; r < OUT; ++r) pred[s * OUT + r] = y[r];
    }
}

// Loss = mean over samples and outputs of (y - t)^2.  Each block handles
// kTrainBlock samples and writes its gradient partial and loss partial.
__global__ __launch_bounds__(kTrainBlock) void k_mlp_fwd_bwd(
    const float *__restrict__ params, const float *__restrict__ feat,
    const float *__restrict__ target, int64_t n, float *__restrict__ gpart,
    float *__restrict__ lpart)
{
    __shared__ float P[NP];
    __shared__ float X[kTrainBlock][IN], H1[kTrainBlock][HID], H2[kTrainBlock][HID];
    __shared__ float DY[kTrainBlock][OUT], D2[kTrainBlock][HID], D1[kTrainBlock][HID];
    __shared__ float LS[kTrainBlock];
    const int t = threadIdx.x;
    for (int i = t; i < NP; i += kTrainBlock) P[i] = params[i];
    __syncthreads();
    const int64_t s = (int64_t)blockIdx.x * kTrainBlock + t;
    const bool ok = s < n;
    float x[IN], h1[HID], h2[HID], y[OUT];
#pragma unroll
    for (int c = 0; c < IN; ++c) x[c] = ok ? feat[s * IN + c] : 0.f;
    mlp_forward_f32(P, x, h1, h2, y);
    const float scale = 2.f / (float)(n * OUT);
    float loss = 0.f, dy[OUT];
#pragma unroll
    for (int r = 0; r < OUT; ++r) {
        const float e = ok ? (y[r] - target[s * OUT + r]) : 0.f;
        loss += e * e;
        dy[r] = e * scale;
    }
    // back through layer 3 and 2
    float d2[HID];
    for (int j = 0; j < HID; ++j) {
        float a = 0.f;
#pragma unroll
        for (int r = 0; r < OUT; ++r) a = fmaf(P[oW3 + r * HID + j], dy[r], a);
        d2[j] = h2[j] > 0.f ? a : 0.f;
    }
    for (int i = 0; i < HID; ++i) {
        float a = 0.f;
        for (int j = 0; j < HID; ++j) a = fmaf(P[oW2 + j * HID + i], d2[j], a);
        D1[t][i] = h1[i] > 0.f ? a : 0.f;
    }
#pragma unroll
    for (int c = 0; c < IN; ++c) X[t][c] = x[c];
    for (int j = 0; j < HID; ++j) {
        H1[t][j] = h1[j];
        H2[t][j] = h2[j];
        D2[t][j] = d2[j];
    }
#pragma unroll
    for (int r = 0; r < OUT; ++r) DY[t][r] = dy[r];
    LS[t] = loss;
    __syncthreads();
    // gradient entries, summed over the block's samples in a fixed order
    for (int p = t; p < NP; p += kTrainBlock) {
        float g = 0.f;
        if (p < oB1) {
            const int r = (p - oW1) / IN, c = (p - oW1) % IN;
            for (int k = 0; k < kTrainBlock; ++k) g = fmaf(D1[k][r], X[k][c], g);
        } else if (p < oW2) {
            const int r = p - oB1;
            for (int k = 0; k < kTrainBlock; ++k) g += D1[k][r];
        } else if (p < oB2) {
            const int r = (p - oW2) / HID, c = (p - oW2) % HID;
            for (int k = 0; k < kTrainBlock; ++k) g = fmaf(D2[k][r], H1[k][c], g);
        } else if (p < oW3) {
            const int r = p - oB2;
            for (int k = 0; k < kTrainBlock; ++k) g += D2[k][r];
        } else if (p < oB3) {
            const int r = (p - oW3) / HID, c = (p - oW3) % HID;
            for (int k = 0; k < kTrainBlock; ++k) g = fmaf(DY[k][r], H2[k][c], g);
        } else {
            const int r = p - oB3;
            for (int k = 0; k < kTrainBlock; ++k) g += DY[k][r];
        }
        gpart[(int64_t)blockIdx.x * NP + p] = g;
    }
    if (t == 0) {
        float l = 0.f;
        for (int k = 0; k < kTrainBlock; ++k) l += LS[k];
        lpart[blockIdx.x] = l;
    }
}

__global__ void k_mlp_reduce_adam(const float *__restrict__ gpart, const float *__restrict__ lpart,
                                  int32_t nblocks, int64_t n, float *__restrict__ params,
                                  float *__restrict__ m, float *__restrict__ v, int32_t step,
                                  float lr, float *__restrict__ grads, float *__restrict__ loss)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < NP) {
        float g = 0.f;
        for (int b = 0; b < nblocks; ++b) g += gpart[(int64_t)b * NP + p];
        if (grads) grads[p] = g;
        if (m && v) {
            const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
            const float mm = b1 * m[p] + (1.f - b1) * g;
            const float vv = b2 * v[p] + (1.f - b2) * g * g;
            m[p] = mm;
            v[p] = vv;
            const float mh = mm / (1.f - powf(b1, (float)step));
            const float vh = vv / (1.f - powf(b2, (float)step));
            params[p] -= lr * mh / (sqrtf(vh) + eps);
        } else {
            params[p] -= lr * g;
        }
    }
    if (p == 0 && loss) {
        float l = 0.f;
        for (int b = 0; b < nblocks; ++b) l += lpart[b];
        *loss = l / (float)(n * OUT);
    }
}

// ------------------------------------------------------- BF16 inference --

__device__ __host__ inline uint16_t f32_to_bf16(float f)
{
    uint32_t u;
    memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
    return (uint16_t)(u >> 16);
}

__device__ inline float bf16_to_f32(uint16_t b)
{
    const uint32_t u = (uint32_t)b << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}

__device__ inline float bf16_round(float f) { return bf16_to_f32(f32_to_bf16(f)); }

// Features of the cell (i, j) of log2 buckets (bucket k >= 1 holds
// [2^(k-1), 2^k)); bucket centres are used.
__device__ inline void cell_features(int i, int j, int small_max, int mid_max, float x[IN])
{
    const float fl = i == 0 ? 0.f : (float)i - 0.5f;
    const float fp = j == 0 ? 0.f : (float)j - 0.5f;
    const float L = exp2f(fl);
    const int bin = L <= (float)small_max ? 0 : (L <= (float)mid_max ? 1 : 2);
    x[0] = fl / 16.f;
    x[1] = fp / 16.f;
    x[2] = (fp - fl) / 16.f;
    x[3] = (float)bin / 2.f;
}

__global__ void k_model_table(const uint16_t *__restrict__ wb, int small_max, int mid_max,
                              uint8_t *__restrict__ table)
{
    __shared__ float P[NP];
    for (int i = threadIdx.x; i < NP; i += blockDim.x) P[i] = bf16_to_f32(wb[i]);
    __syncthreads();
    const int cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= kTableDim * kTableDim) return;
    const int i = cell / kTableDim, j = cell % kTableDim;
    float x[IN], h1[HID], h2[HID], y[OUT];
    cell_features(i, j, small_max, mid_max, x);
#pragma unroll
    for (int c = 0; c < IN; ++c) x[c] = bf16_round(x[c]);
    for (int r = 0; r < HID; ++r) {
        float a = P[oB1 + r];
        for (int c = 0; c < IN; ++c) a = fmaf(P[oW1 + r * IN + c], x[c], a);
        h1[r] = bf16_round(a > 0.f ? a : 0.f);
    }
    for (int r = 0; r < HID; ++r) {
        float a = P[oB2 + r];
        for (int c = 0; c < HID; ++c) a = fmaf(P[oW2 + r * HID + c], h1[c], a);
        h2[r] = bf16_round(a > 0.f ? a : 0.f);
    }
    int best = 0;
    float bv = 0.f;
    for (int r = 0; r < OUT; ++r) {
        float a = P[oB3 + r];
        for (int c = 0; c < HID; ++c) a = fmaf(P[oW3 + r * HID + c], h2[c], a);
        y[r] = a;
        if (r == 0 || a < bv) {
            bv = a;
            best = r;
        }
    }
    // a cell without an owner or task list never occurs; keep the fixed choice there
    table[cell] = (uint8_t)best;
}

// --------------------------------------------------------------- timing --

// Samples entries of the oriented task array: owner (rank), task vertex
// (rank) and the start of the task's useful suffix of N+(task).
__global__ void k_sample_tasks(const int64_t *__restrict__ toff, const VInfo *__restrict__ vi,
                               const TaskRec *__restrict__ tasks, const int64_t *__restrict__ pcoff,
                               int32_t nv, int64_t ntasks,
                               int64_t ncand, uint64_t seed, int32_t lmax,
                               int32_t *__restrict__ so, int32_t *__restrict__ st,
                               int32_t *__restrict__ sk)
{
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < ncand;
         s += (int64_t)gridDim.x * blockDim.x) {
        uint64_t z = seed + 0x9E3779B97F4A7C15ULL * (uint64_t)(s + 1);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        z ^= z >> 31;
        const int64_t e = (int64_t)((z >> 11) % (uint64_t)ntasks);
        int32_t lo = 0, hi = nv - 1;  // owner: last row whose first task is <= e
        while (lo < hi) {
            const int32_t mid = lo + ((hi - lo + 1) >> 1);
            if (toff[mid] <= e) lo = mid; else hi = mid - 1;
        }
        const VInfo vb = vi[lo];
        int32_t own = -1, task = -1, kp = 0;
        const int32_t L = vb.len - vb.dm;
        if (L >= 1 && L <= lmax) {
            const TaskRec tk = tasks[e];
            // the task vertex: the plus-array row holding entry base - 1
            const int64_t x = tk.base() - 1;
            int32_t a = 0, ah = nv - 1;
            while (a < ah) {
                const int32_t mid = a + ((ah - a + 1) >> 1);
                if (pcoff[mid] <= x) a = mid; else ah = mid - 1;
            }
            const int32_t p = (vi[a].len - vi[a].dm) - tk.len() - 1;  // position of the owner in N+(a)
            if (tk.len() >= 1) {
                own = lo;
                task = a;
                kp = p;
            }
        }
        so[s] = own;
        st[s] = task;
        sk[s] = kp;
    }
}

struct Tbl {
    int4 *HB;
    int nbb;
    __device__ uint32_t b1(int32_t w) const { return nbb ? (((uint32_t)w * 0x9E3779B1u) >> (32 - nbb)) : 0u; }
    __device__ uint32_t b2(int32_t w) const
    {
        return nbb ? ((((uint32_t)w ^ 0x2C1B3C6Du) * 0x297A2D39u) >> (32 - nbb)) : 0u;
    }
    __device__ void insert(int32_t w) const
    {
        const uint32_t m = (1u << nbb) - 1;
        const uint32_t x = b1(w), y = b2(w);
        const int4 vx = HB[x], vy = HB[y];
        const int fx = (vx.x >= 0) + (vx.y >= 0) + (vx.z >= 0) + (vx.w >= 0);
        const int fy = (vy.x >= 0) + (vy.y >= 0) + (vy.z >= 0) + (vy.w >= 0);
        const uint32_t order[2] = {fx <= fy ? x : y, fx <= fy ? y : x};
        for (int q = 0; q < 2; ++q) {
            int *s = reinterpret_cast<int *>(&HB[order[q]]);
            for (int k = 0; k < 4; ++k)
                if (atomicCAS(&s[k], -1, w) == -1) return;
        }
        for (uint32_t b = (x + 1) & m;; b = (b + 1) & m) {
            int *s = reinterpret_cast<int *>(&HB[b]);
            for (int k = 0; k < 4; ++k)
                if (atomicCAS(&s[k], -1, w) == -1) return;
        }
    }
    __device__ bool has(int32_t w) const
    {
        const uint32_t x = b1(w), y = b2(w);
        const int4 vx = HB[x], vy = HB[y];
        bool f = vx.x == w || vx.y == w || vx.z == w || vx.w == w || vy.x == w || vy.y == w ||
                 vy.z == w || vy.w == w;
        if (!f && vx.w >= 0 && vy.w >= 0) {
            const uint32_t m = (1u << nbb) - 1;
            for (uint32_t b = (x + 1) & m;; b = (b + 1) & m) {
                const int4 v = HB[b];
                if (v.x == w || v.y == w || v.z == w || v.w == w) {
                    f = true;
                    break;
                }
                if (v.w < 0) break;
            }
        }
        return f;
    }
};

constexpr int kTimeLmax = 4096;

// One wavefront per sample: the owner's N+ list goes into an LDS table, then
// the three strategies run back to back, each bracketed by the shader clock.
__global__ __launch_bounds__(64) void k_time_tasks(const int32_t *__restrict__ so,
                                                   const int32_t *__restrict__ st,
                                                   const int32_t *__restrict__ sk, int64_t n,
                                                   const VInfo *__restrict__ vi,
                                                   const int32_t *__restrict__ padj,
                                                   float *__restrict__ cost,
                                                   int32_t *__restrict__ agree)
{
    __shared__ int4 HB[kTimeLmax / 2];
    const int lane = threadIdx.x;
    const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
    (void)lt;
    for (int64_t s = blockIdx.x; s < n; s += gridDim.x) {
        const int32_t b = so[s], a = st[s], kp = sk[s];
        const VInfo vb = vi[b], va = vi[a];
        const int32_t L = vb.len - vb.dm, pl = va.len - va.dm - kp - 1;
        const int32_t *S = padj + vb.rb;
        const int32_t *T = padj + va.rb + kp + 1;
        int nbb = 0;
        while ((1 << nbb) * 2 < L) ++nbb;
        const Tbl tb{HB, nbb};
        for (int i = lane; i < (1 << nbb); i += 64) HB[i] = make_int4(-1, -1, -1, -1);
        __syncthreads();
        for (int i = lane; i < L; i += 64) tb.insert(S[i]);
        __syncthreads();
        // hash: stream the task list through the table
        uint64_t t0 = clock64();
        uint32_t hh = 0;
        for (int32_t k = lane; k < pl; k += 64) hh += tb.has(T[k]) ? 1u : 0u;
        for (int d = 32; d > 0; d >>= 1) hh += __shfl_xor(hh, d);
        __syncthreads();
        uint64_t t1 = clock64();
        // binary search: each owner entry looked up in the task list
        uint32_t hb = 0;
        for (int32_t k = lane; k < L; k += 64) {
            const int32_t w = S[k];
            int32_t lo = 0, hi = pl;
            while (lo < hi) {
                const int32_t mid = (lo + hi) >> 1;
                if (T[mid] < w) lo = mid + 1; else hi = mid;
            }
            hb += (lo < pl && T[lo] == w) ? 1u : 0u;
        }
        for (int d = 32; d > 0; d >>= 1) hb += __shfl_xor(hb, d);
        __syncthreads();
        uint64_t t2 = clock64();
        // merge: one lane walks both lists
        uint32_t hm = 0;
        if (lane == 0) {
            int32_t i = 0, k = 0;
            while (i < pl && k < L) {
                const int32_t x = T[i], y = S[k];
                hm += x == y;
                i += x <= y;
                k += x >= y;
            }
        }
        hm = __shfl(hm, 0);
        __syncthreads();
        uint64_t t3 = clock64();
        if (lane == 0) {
            cost[s * 3 + 0] = (float)(t3 - t2);
            cost[s * 3 + 1] = (float)(t2 - t1);
            cost[s * 3 + 2] = (float)(t1 - t0);
            agree[s] = (hh == hb && hb == hm) ? 1 : 0;
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------- state ----

struct LearnState {
    uint16_t *wb = nullptr;   // BF16 weights (device)
    uint8_t *table = nullptr; // strategy table (device)
    bool ready = false;
    tc_schedule tuned = {};
    bool has_tuned = false;
};
LearnState g_learn;
std::mutex g_learn_mu;

}  // namespace

const uint8_t *learned_table_device(int32_t *dim)
{
    std::lock_guard<std::mutex> lk(g_learn_mu);
    *dim = kTableDim;
    return g_learn.ready ? g_learn.table : nullptr;
}

bool schedule_override(tc_schedule *s)
{
    std::lock_guard<std::mutex> lk(g_learn_mu);
    if (!g_learn.has_tuned) return false;
    const int32_t method = s->method, policy = s->policy;
    *s = g_learn.tuned;
    s->method = method;
    s->policy = policy;
    return true;
}

int collect_timings(Graph &g, int64_t max_samples, uint64_t seed, float *feat, float *cost,
                    int64_t *n_out)
{
    *n_out = 0;
    if (g.symmetric < 0 || g.self_loops < 0) TC_TRY(validate_graph(g, true, nullptr));
    if (g.symmetric != 1 || g.self_loops != 0) {
        set_last_error(TC_ERR_INVALID, "timing collection needs a symmetric loop-free graph");
        return TC_ERR_INVALID;
    }
    if (g.st.nnz == 0) return TC_OK;
    TC_TRY(prepare_views(g, true));
    if (g.ntasks == 0) return TC_OK;
    hipStream_t s = g.stream;
    const int64_t ncand = 4 * max_samples;
    DevBuf so, st, sk, cst, agr;
    auto done = [&](int r) {
        so.release();
        st.release();
        sk.release();
        cst.release();
        agr.release();
        return r;
    };
    int r;
    if ((r = so.ensure((size_t)ncand * 4)) || (r = st.ensure((size_t)ncand * 4)) ||
        (r = sk.ensure((size_t)ncand * 4)))
        return done(r);
    hipLaunchKernelGGL(k_sample_tasks, dim3(grid_for(ncand, 256, 4096)), dim3(256), 0, s,
                       g.coff.as<int64_t>(), g.vinfo.as<VInfo>(), g.tasks.as<TaskRec>(),
                       g.ov[6].as<int64_t>(), g.st.nv,
                       g.ntasks, ncand, seed, kTimeLmax, so.as<int32_t>(), st.as<int32_t>(),
                       sk.as<int32_t>());
    std::vector<int32_t> ho(ncand), ht(ncand), hk(ncand);
    if (hipMemcpy(ho.data(), so.p, (size_t)ncand * 4, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(ht.data(), st.p, (size_t)ncand * 4, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(hk.data(), sk.p, (size_t)ncand * 4, hipMemcpyDeviceToHost) != hipSuccess) {
        set_last_error(TC_ERR_HIP, "collect_timings: sample readback");
        return done(TC_ERR_HIP);
    }
    int64_t n = 0;
    for (int64_t i = 0; i < ncand && n < max_samples; ++i)
        if (ho[i] >= 0) {
            ho[n] = ho[i];
            ht[n] = ht[i];
            hk[n] = hk[i];
            ++n;
        }
    if (n == 0) return done(TC_OK);
    if ((r = cst.ensure((size_t)n * 3 * sizeof(float))) || (r = agr.ensure((size_t)n * 4)))
        return done(r);
    if (hipMemcpy(so.p, ho.data(), (size_t)n * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(st.p, ht.data(), (size_t)n * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(sk.p, hk.data(), (size_t)n * 4, hipMemcpyHostToDevice) != hipSuccess) {
        set_last_error(TC_ERR_HIP, "collect_timings: upload");
        return done(TC_ERR_HIP);
    }
    int dev = 0, cus = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    hipLaunchKernelGGL(k_time_tasks, dim3((unsigned)std::min<int64_t>(n, (int64_t)cus * 4)), dim3(64),
                       0, s, so.as<int32_t>(), st.as<int32_t>(), sk.as<int32_t>(), n,
                       g.vinfo.as<VInfo>(),
                       g.padj.as<int32_t>(), cst.as<float>(), agr.as<int32_t>());
    std::vector<float> hc((size_t)n * 3);
    std::vector<int32_t> ha((size_t)n);
    std::vector<VInfo> hv((size_t)g.st.nv);
    if (hipStreamSynchronize(s) != hipSuccess ||
        hipMemcpy(hc.data(), cst.p, hc.size() * 4, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(ha.data(), agr.p, ha.size() * 4, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(hv.data(), g.vinfo.p, hv.size() * sizeof(VInfo), hipMemcpyDeviceToHost) !=
            hipSuccess) {
        set_last_error(TC_ERR_HIP, "collect_timings: kernels");
        return done(TC_ERR_HIP);
    }
    for (int64_t i = 0; i < n; ++i)
        if (!ha[i]) {
            set_last_error(TC_ERR_HIP, "collect_timings: strategies disagree on a sampled task");
            return done(TC_ERR_HIP);
        }
    const tc_schedule d = default_schedule();
    for (int64_t i = 0; i < n; ++i) {
        const VInfo vb = hv[ho[i]], va = hv[ht[i]];
        const float L = (float)(vb.len - vb.dm), pl = (float)(va.len - va.dm - hk[i] - 1);
        const float fl = std::log2(L), fp = std::log2(pl);
        const int bin = L <= d.small_max ? 0 : (L <= d.mid_max ? 1 : 2);
        feat[i * 4 + 0] = fl / 16.f;
        feat[i * 4 + 1] = fp / 16.f;
        feat[i * 4 + 2] = (fp - fl) / 16.f;
        feat[i * 4 + 3] = (float)bin / 2.f;
        for (int q = 0; q < 3; ++q) cost[i * 3 + q] = hc[i * 3 + q];
    }
    *n_out = n;
    return done(TC_OK);
}

int mlp_train_step(float *params, float *m, float *v, int32_t step, const float *feat,
                   const float *target, int64_t n, float lr, float *grads, float *loss)
{
    const int32_t nb = (int32_t)ceil_div(n, kTrainBlock);
    DevBuf gp, lp, dl;
    auto done = [&](int r) {
        gp.release();
        lp.release();
        dl.release();
        return r;
    };
    int r;
    if ((r = gp.ensure((size_t)nb * NP * sizeof(float))) || (r = lp.ensure((size_t)nb * sizeof(float))) ||
        (r = dl.ensure(sizeof(float))))
        return done(r);
    hipLaunchKernelGGL(k_mlp_fwd_bwd, dim3(nb), dim3(kTrainBlock), 0, nullptr, params, feat, target, n,
                       gp.as<float>(), lp.as<float>());
    hipLaunchKernelGGL(k_mlp_reduce_adam, dim3((NP + 255) / 256), dim3(256), 0, nullptr,
                       gp.as<float>(), lp.as<float>(), nb, n, params, m, v, step, lr, grads,
                       dl.as<float>());
    float h = 0.f;
    if (hipGetLastError() != hipSuccess ||
        hipMemcpy(&h, dl.p, sizeof(float), hipMemcpyDeviceToHost) != hipSuccess) {
        set_last_error(TC_ERR_HIP, "mlp_train_step: kernels");
        return done(TC_ERR_HIP);
    }
    if (loss) *loss = h;
    return done(TC_OK);
}

int mlp_forward(const float *params, const float *feat, int64_t n, float *pred)
{
    if (n == 0) return TC_OK;
    hipLaunchKernelGGL(k_mlp_forward, dim3(grid_for(n, 256, 4096)), dim3(256), 0, nullptr, params,
                       feat, n, pred);
    TC_HIP(hipGetLastError());
    TC_HIP(hipDeviceSynchronize());
    return TC_OK;
}

int model_load(const float *params_host)
{
    std::lock_guard<std::mutex> lk(g_learn_mu);
    if (!params_host) {  // unload: back to the fixed heuristic
        g_learn.ready = false;
        return TC_OK;
    }
    for (int i = 0; i < NP; ++i)
        if (!std::isfinite(params_host[i])) {
            set_last_error(TC_ERR_INVALID, "tc_model_load: non-finite parameter");
            return TC_ERR_INVALID;
        }
    if (!g_learn.wb) {
        TC_HIP(hipMalloc(&g_learn.wb, NP * sizeof(uint16_t)));
        TC_HIP(hipMalloc(&g_learn.table, kTableDim * kTableDim));
    }
    std::vector<uint16_t> hb(NP);
    for (int i = 0; i < NP; ++i) hb[i] = f32_to_bf16(params_host[i]);
    TC_HIP(hipMemcpy(g_learn.wb, hb.data(), NP * sizeof(uint16_t), hipMemcpyHostToDevice));
    const tc_schedule d = default_schedule();
    hipLaunchKernelGGL(k_model_table, dim3((kTableDim * kTableDim + 255) / 256), dim3(256), 0, nullptr,
                       g_learn.wb, d.small_max, d.mid_max, g_learn.table);
    TC_HIP(hipGetLastError());
    TC_HIP(hipDeviceSynchronize());
    g_learn.ready = true;
    return TC_OK;
}

// A strategy table given directly (an in-situ sweep's choices, or a table a
// caller saved): entries are strategies 0..3 (merge, bsearch, hash, merge
// path), read by POLICY_LEARNED like a model's table.
int strategy_table_set(const uint8_t *table, int32_t dim)
{
    std::lock_guard<std::mutex> lk(g_learn_mu);
    if (!table) {
        g_learn.ready = false;
        return TC_OK;
    }
    if (dim != kTableDim) {
        set_last_error(TC_ERR_INVALID, "tc_strategy_table_set: the table must be 32 x 32");
        return TC_ERR_INVALID;
    }
    for (int i = 0; i < kTableDim * kTableDim; ++i)
        if (table[i] > 3) {
            set_last_error(TC_ERR_INVALID, "tc_strategy_table_set: strategies are 0..3");
            return TC_ERR_INVALID;
        }
    if (!g_learn.wb) {
        TC_HIP(hipMalloc(&g_learn.wb, NP * sizeof(uint16_t)));
        TC_HIP(hipMalloc(&g_learn.table, kTableDim * kTableDim));
    }
    TC_HIP(hipMemcpy(g_learn.table, table, kTableDim * kTableDim, hipMemcpyHostToDevice));
    g_learn.ready = true;
    return TC_OK;
}

int model_table(uint8_t *table, int32_t *dim)
{
    std::lock_guard<std::mutex> lk(g_learn_mu);
    *dim = kTableDim;
    if (!table) return TC_OK;
    if (!g_learn.ready) {
        set_last_error(TC_ERR_INVALID, "tc_model_table: no model loaded");
        return TC_ERR_INVALID;
    }
    TC_HIP(hipMemcpy(table, g_learn.table, kTableDim * kTableDim, hipMemcpyDeviceToHost));
    return TC_OK;
}

int set_default_schedule(const tc_schedule *s)
{
    std::lock_guard<std::mutex> lk(g_learn_mu);
    g_learn.has_tuned = s != nullptr;
    if (s) g_learn.tuned = *s;
    return TC_OK;
}

int autotune(Graph &g, tc_schedule *best, double *best_ms)
{
    tc_schedule base = default_schedule();
    const int32_t smalls[] = {32, 64};
    const int32_t mids[] = {256, 512};
    const int32_t flushes[] = {256, 1024, 4096};
    const int32_t jobs[] = {4096, 16384};
    double bms = 1e300;
    tc_schedule bs = base;
    uint64_t ref = 0;
    bool have_ref = false;
    for (int32_t sm : smalls)
        for (int32_t md : mids)
            for (int32_t fl : flushes)
                for (int32_t jb : jobs) {
                    tc_schedule s = base;
                    s.small_max = sm;
                    s.mid_max = md;
                    s.flush_items = fl;
                    s.job_items = jb;
                    double ms = 1e300;
                    for (int rep = 0; rep < 2; ++rep) {
                        tc_stats st = {};
                        uint64_t glob = 0;
                        TC_TRY(static_count(g, nullptr, &glob, s, &st, false));
                        if (!have_ref) {
                            ref = glob;
                            have_ref = true;
                        } else if (glob != ref) {
                            set_last_error(TC_ERR_HIP, "autotune: counts differ between schedules");
                            return TC_ERR_HIP;
                        }
                        ms = std::min(ms, st.ms_total);
                    }
                    if (ms < bms) {
                        bms = ms;
                        bs = s;
                    }
                }
    {
        std::lock_guard<std::mutex> lk(g_learn_mu);
        g_learn.tuned = bs;
        g_learn.has_tuned = true;
    }
    *best = bs;
    if (best_ms) *best_ms = bms;
    return TC_OK;
}

}  // namespace tc
