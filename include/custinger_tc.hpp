// custinger_tc.hpp -- C++ host surface of the MI355X triangle-counting engine.
//
// Header-only, over the C-ABI in custinger_tc.h (plain C types; no HIP or
// torch types).  It keeps the names and argument meaning of the call sites the
// reference's triangle-counting drivers use, so code written against
// cuStingerAlg's static/streaming triangle counting compiles with small edits:
//
//   typedefs                 vertexId_t, length_t (int32), triangle_t (int32)
//                            (include/static_triangle_counting/cct.hpp:12)
//   cuStinger / cuStingerInitConfig / initializeCuStinger / freecuStinger
//                            (tests/cct.cpp:261-281, tests/cct_upd.cpp:227-244)
//   getNumberEdgesUsed / getNumberEdgesAllocated (tests/cct_upd.cpp:127-128)
//   BatchUpdateData / BatchUpdate / sortDeviceBUD / edgeInsertionsSorted /
//   edgeDeletions            (tests/cct_upd.cpp:285,391-398; tests/katz-tester.cu:244)
//   callDeviceAllTriangles   (include/static_triangle_counting/cct.hpp:15-17)
//   callDeviceNewTriangles   (src/streaming_triangle_counting/cct_upd.cu:412-415)
//   callDeviceAllTrianglesCSR(tests/cct_csr.cu:279-285)
//   cuStingerAlgs::StaticAlgorithm (include/algs.cuh:20-29) with
//   StaticTriangleCounting, and StreamingTriangleCounting in the shape of
//   katzCentralityStreaming (include/streaming_katz_centrality/katz.cuh:21-58).
//
// Errors: the C-ABI returns status codes; this layer throws cuStingerError
// (derived from std::runtime_error) with the library's message.  The
// reference's free functions return void, so they throw as well.
#ifndef CUSTINGER_TC_HPP
#define CUSTINGER_TC_HPP

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <climits>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "custinger_tc.h"

typedef int32_t vertexId_t;
typedef int32_t length_t;
typedef int32_t triangle_t;

class cuStingerError : public std::runtime_error {
public:
    cuStingerError(int status, const std::string &what) : std::runtime_error(what), status(status) {}
    int status;
};

inline void tcCheck(int status)
{
    if (status != TC_OK) throw cuStingerError(status, std::string(tc_status_string(status)) + ": " +
                                                          tc_last_error());
}

// ------------------------------------------------------- device helpers --
inline void *allocDeviceArray(length_t elements, int32_t elementSize)
{
    void *p = nullptr;
    tcCheck(tc_device_alloc(&p, (uint64_t)elements * (uint64_t)elementSize));
    return p;
}
inline void freeDeviceArray(void *p) { tcCheck(tc_device_free(p)); }
inline void copyArrayHostToDevice(const void *h, void *d, length_t elements, int32_t elementSize)
{
    tcCheck(tc_memcpy(d, h, (uint64_t)elements * (uint64_t)elementSize, 0));
}
inline void copyArrayDeviceToHost(const void *d, void *h, length_t elements, int32_t elementSize)
{
    tcCheck(tc_memcpy(h, d, (uint64_t)elements * (uint64_t)elementSize, 1));
}
inline void copyArrayDeviceToDevice(const void *s, void *d, length_t elements, int32_t elementSize)
{
    tcCheck(tc_memcpy(d, s, (uint64_t)elements * (uint64_t)elementSize, 2));
}

// ------------------------------------------------------------- the store --
// Block-capacity policy: the reference passes allocator callbacks; here an
// allocator maps a degree to the capacity of the vertex's block.
typedef length_t (*cuStingerAllocator)(length_t elements);
inline length_t defaultInitAllocater(length_t elements) { return elements; }
inline length_t defaultUpdateAllocater(length_t elements) { return elements * 2; }

enum cuStingerInitState { eInitStateEmpty = 0, eInitStateCSR = 1, eInitStateEdgeList = 2 };

struct cuStingerInitConfig {
    cuStingerInitState initState = eInitStateCSR;
    length_t maxNV = 0;
    bool useVWeight = false;
    bool isSemantic = false;
    bool useEWeight = false;
    length_t csrNV = 0;
    length_t csrNE = 0;
    const length_t *csrOff = nullptr;
    const vertexId_t *csrAdj = nullptr;
    const void *csrVW = nullptr;
    const void *csrEW = nullptr;
};

class BatchUpdate;

class cuStinger {
public:
    cuStinger(cuStingerAllocator initAlloc = defaultInitAllocater,
              cuStingerAllocator updateAlloc = defaultUpdateAllocater)
        : initAlloc_(initAlloc), updateAlloc_(updateAlloc)
    {
    }
    ~cuStinger() { freecuStinger(); }
    cuStinger(const cuStinger &) = delete;
    cuStinger &operator=(const cuStinger &) = delete;

    void initializeCuStinger(const cuStingerInitConfig &cfg)
    {
        if (cfg.initState != eInitStateCSR && cfg.initState != eInitStateEmpty)
            throw cuStingerError(TC_ERR_UNSUPPORTED, "only CSR and empty initialisation");
        if (cfg.useVWeight || cfg.useEWeight || cfg.isSemantic)
            throw cuStingerError(TC_ERR_UNSUPPORTED, "vertex/edge weights and types are not stored");
        freecuStinger();
        nv = cfg.initState == eInitStateCSR ? cfg.csrNV : cfg.maxNV;
        std::vector<int64_t> off((size_t)nv + 1, 0);
        if (cfg.initState == eInitStateCSR)
            for (length_t v = 0; v <= nv; ++v) off[(size_t)v] = cfg.csrOff[v];
        const int64_t nnz = off[(size_t)nv];
        // capacity headroom from the init allocator, sampled at a large degree
        const int64_t probe = 1 << 20;
        int32_t slack = (int32_t)((int64_t)initAlloc_((length_t)probe) * 100 / probe - 100);
        if (slack < 0) slack = 0;
        tcCheck(tc_graph_create_csr(nv, nnz, off.data(), nnz ? cfg.csrAdj : nullptr, TC_HOST, 0, slack,
                                    &g_));
        // growth of rows that outgrow their blocks: the update allocator,
        // read as the line a*x + b through two sample sizes
        const double x1 = 1024.0, x2 = 1048576.0;
        const double f1 = (double)updateAlloc_((length_t)x1), f2 = (double)updateAlloc_((length_t)x2);
        const double a = (f2 - f1) / (x2 - x1);
        int32_t mul_x16 = (int32_t)(a * 16.0 + 0.5);
        int32_t add = (int32_t)std::max(0.0, f1 - a * x1 + 0.5);
        if (mul_x16 < 16 && add == 0) mul_x16 = 16;  // never below the used length
        tcCheck(tc_graph_set_growth(g_, mul_x16, add));
    }
    void freecuStinger()
    {
        if (g_) tc_graph_destroy(g_);
        g_ = nullptr;
    }

    length_t getMaxNV() const { return nv; }
    int64_t getNumberEdgesUsed() const { return info().nnz; }
    int64_t getNumberEdgesAllocated() const { return info().allocated; }

    // Inserts the batch's directed entries as given (allocs = blocks moved to
    // grow); the per-vertex duplicate counts of its BatchUpdateData are filled.
    void edgeInsertionsSorted(BatchUpdate &bu, length_t &allocs);
    void edgeInsertions(BatchUpdate &bu, length_t &allocs) { edgeInsertionsSorted(bu, allocs); }
    // Removes the batch's directed entries (tests/katz-tester.cu:244).
    void edgeDeletions(BatchUpdate &bu);

    tc_graph_info info() const
    {
        tc_graph_info i;
        tcCheck(tc_graph_get_info(g_, &i));
        return i;
    }
    tc_graph *handle() const { return g_; }

    length_t nv = 0;

private:
    tc_graph *g_ = nullptr;
    cuStingerAllocator initAlloc_, updateAlloc_;
};

// Host-side batch of directed entries (the reference fills both directions
// for an undirected graph, tests/cct_upd.cpp:337-341; katz-tester's directed
// batches one, tests/katz-tester.cu:211-222).  getvNumDuplicates()[v] counts
// the entries with source v that edgeInsertions found already present or
// repeated in the batch (tests/cct_upd.cpp:335 zeroes it before the update).
class BatchUpdateData {
public:
    BatchUpdateData(length_t batchSize, bool isHost = true, length_t numVertices = 0)
        : src_((size_t)batchSize), dst_((size_t)batchSize), dups_((size_t)numVertices, 0),
          nv_(numVertices)
    {
        (void)isHost;
    }
    vertexId_t *getSrc() { return src_.data(); }
    vertexId_t *getDst() { return dst_.data(); }
    length_t *getBatchSize() { return &size_; }
    length_t *getNumVertices() { return &nv_; }
    length_t *getvNumDuplicates() { return dups_.data(); }
    length_t size() const { return (length_t)src_.size(); }

private:
    std::vector<vertexId_t> src_, dst_;
    std::vector<length_t> dups_;
    length_t nv_ = 0;
    length_t size_ = (length_t)src_.size();
};

// The batch's device arrays in the shape of the reference's device-side
// BatchUpdateData (getDeviceBUD()->..., include/operators.cuh:334-344): the
// getters return device pointers; devicePtr() returns this host view.
struct BatchUpdateDeviceData {
    vertexId_t *src = nullptr, *dst = nullptr;
    length_t size = 0;
    vertexId_t *getSrc() const { return src; }
    vertexId_t *getDst() const { return dst; }
    const length_t *getBatchSize() const { return &size; }
    const BatchUpdateDeviceData *devicePtr() const { return this; }
};

// Device copy of a batch plus its sorted form and duplicate flags.
class BatchUpdate {
public:
    explicit BatchUpdate(BatchUpdateData &bud) : nb_(bud.size()), bud_(&bud)
    {
        src_ = alloc<int32_t>(nb_);
        dst_ = alloc<int32_t>(nb_);
        copyArrayHostToDevice(bud.getSrc(), src_, nb_, 4);
        copyArrayHostToDevice(bud.getDst(), dst_, nb_, 4);
    }
    ~BatchUpdate()
    {
        for (void *p : {(void *)src_, (void *)dst_, (void *)ssrc_, (void *)sdst_, (void *)dup_,
                        (void *)boff_})
            if (p) tc_device_free(p);
    }
    BatchUpdate(const BatchUpdate &) = delete;
    BatchUpdate &operator=(const BatchUpdate &) = delete;

    // The reference sorts here and flags duplicates when the batch is
    // inserted; the sort is done together with the flags (which need the
    // pre-insert store) by edgeInsertionsSorted.  blockdim is accepted for
    // signature parity.
    void sortDeviceBUD(const int blockdim = 0)
    {
        (void)blockdim;
        sort_requested_ = true;
    }
    // Sorted entries, duplicate flags (repeats and entries already in the
    // store) and batch offsets against the store's current contents.
    void prepare(cuStinger &custing);
    length_t getBatchSize() const { return nb_; }
    BatchUpdateData *getHostBUD() { return bud_; }
    BatchUpdateDeviceData *getDeviceBUD()
    {
        dview_.src = src_;
        dview_.dst = dst_;
        dview_.size = nb_;
        return &dview_;
    }
    bool prepared() const { return prepared_; }
    bool sortRequested() const { return sort_requested_; }
    const int32_t *deviceSrc() const { return src_; }
    const int32_t *deviceDst() const { return dst_; }
    const int32_t *sortedSrc() const { return ssrc_; }
    const int32_t *sortedDst() const { return sdst_; }
    const int32_t *duplicateFlags() const { return dup_; }
    const int64_t *batchOffsets() const { return boff_; }

private:
    template <class T> static T *alloc(int64_t n)
    {
        void *p = nullptr;
        tcCheck(tc_device_alloc(&p, (uint64_t)(n > 0 ? n : 1) * sizeof(T)));
        return static_cast<T *>(p);
    }
    length_t nb_;
    BatchUpdateData *bud_;
    BatchUpdateDeviceData dview_;
    int32_t *src_ = nullptr, *dst_ = nullptr, *ssrc_ = nullptr, *sdst_ = nullptr, *dup_ = nullptr;
    int64_t *boff_ = nullptr;
    bool prepared_ = false;
    bool sort_requested_ = false;
    friend class cuStinger;
};

inline void BatchUpdate::prepare(cuStinger &custing)
{
    const length_t nv = custing.getMaxNV();
    if (!ssrc_) {
        ssrc_ = alloc<int32_t>(nb_);
        sdst_ = alloc<int32_t>(nb_);
        dup_ = alloc<int32_t>(nb_);
        boff_ = alloc<int64_t>(nv + 1);
    }
    tcCheck(tc_batch_prepare(custing.handle(), nb_, src_, dst_, ssrc_, sdst_, dup_, boff_));
    prepared_ = true;
}

inline void cuStinger::edgeInsertionsSorted(BatchUpdate &bu, length_t &allocs)
{
    bu.prepare(*this);  // flags against the pre-insert store, as the delta passes need
    const int64_t nb = bu.getBatchSize();
    // per-vertex duplicates: entries the insertion leaves alone, by source
    BatchUpdateData *bud = bu.getHostBUD();
    const length_t nvd = bud ? *bud->getNumVertices() : 0;
    if (nb > 0 && nvd > 0) {
        std::vector<int32_t> ssrc((size_t)nb), dup((size_t)nb);
        copyArrayDeviceToHost(bu.sortedSrc(), ssrc.data(), nb, 4);
        copyArrayDeviceToHost(bu.duplicateFlags(), dup.data(), nb, 4);
        length_t *nd = bud->getvNumDuplicates();
        for (int64_t i = 0; i < nb; ++i)
            if (dup[(size_t)i] && ssrc[(size_t)i] < nvd) ++nd[ssrc[(size_t)i]];
    }
    // the entries as given: both directions of an undirected batch, one of a
    // directed one
    tc_batch_stats st;
    tcCheck(tc_batch_update_directed(g_, bu.deviceSrc(), bu.deviceDst(), nullptr, nb, &st));
    allocs = (length_t)st.reallocations;
}

inline void cuStinger::edgeDeletions(BatchUpdate &bu)
{
    const int64_t nb = bu.getBatchSize();
    std::vector<uint8_t> del((size_t)(nb > 0 ? nb : 1), 1);
    void *d = nullptr;
    tcCheck(tc_device_alloc(&d, (uint64_t)del.size()));
    int st = tc_memcpy(d, del.data(), del.size(), 0);
    tc_batch_stats bs;
    if (st == TC_OK)
        st = tc_batch_update_directed(g_, bu.deviceSrc(), bu.deviceDst(), (const uint8_t *)d, nb, &bs);
    tc_device_free(d);
    tcCheck(st);
}

// compareCUS (src/streaming_triangle_counting/cct_upd.cu:476-484): compares the
// two stores on the device and, like the reference's kernel, prints the
// mismatches (first differing vertex and counts); returns the number of
// differing rows.
inline int64_t compareCUS(cuStinger *cus1, cuStinger *cus2)
{
    int64_t sz = 0, ent = 0;
    int32_t first = -1;
    tcCheck(tc_graph_compare(cus1->handle(), cus2->handle(), &sz, &ent, &first));
    if (sz || ent)
        std::printf("stores differ: %lld rows of different length, %lld rows with different "
                    "entries, first at vertex %d\n", (long long)sz, (long long)ent, first);
    return sz + ent;
}

// ----------------------------------------------------- compat functions --

// callDeviceAllTriangles (cct.hpp:15-17): outPutTriangles is a device array of
// at least nv int32; it is overwritten with 2*t(v) (uint32 bits).  The launch
// parameters of the reference are accepted and ignored: the schedule picks
// its own.
inline void callDeviceAllTriangles(cuStinger &custing, triangle_t *const outPutTriangles,
                                   const int threads_per_block, const int number_blocks,
                                   const int shifter, const int thread_blocks, const int blockdim)
{
    (void)threads_per_block, (void)number_blocks, (void)shifter, (void)thread_blocks, (void)blockdim;
    tcCheck(tc_static_count(custing.handle(), reinterpret_cast<uint32_t *>(outPutTriangles), nullptr,
                            nullptr, nullptr));
}

// callDeviceNewTriangles (cct_upd.cu:412-415): the batch must have been
// sorted (sortDeviceBUD, before insertion) and inserted
// (edgeInsertionsSorted); the three delta passes ACCUMULATE into
// outPutTriangles, as the reference's do.
inline void callDeviceNewTriangles(cuStinger &custing, BatchUpdate &bu,
                                   triangle_t *const outPutTriangles, const int threads_per_block,
                                   const int number_blocks, const int shifter,
                                   const int thread_blocks, const int blockdim,
                                   triangle_t *const h_triangles, triangle_t *const h_triangles_t)
{
    (void)threads_per_block, (void)number_blocks, (void)shifter, (void)thread_blocks, (void)blockdim;
    (void)h_triangles, (void)h_triangles_t;
    if (!bu.prepared())
        throw cuStingerError(TC_ERR_INVALID,
                             "callDeviceNewTriangles: insert the batch with edgeInsertionsSorted first");
    tcCheck(tc_stream_delta_passes(custing.handle(), bu.getBatchSize(), bu.sortedSrc(), bu.sortedDst(),
                                   bu.duplicateFlags(), bu.batchOffsets(),
                                   reinterpret_cast<uint32_t *>(outPutTriangles), 7));
}

// callDeviceAllTrianglesCSR (tests/cct_csr.cu:279-285): device CSR with
// int32 offsets; counts with the CSR kernel's convention (sum = 3T).
inline void callDeviceAllTrianglesCSR(const vertexId_t nv, const length_t *d_off,
                                      const vertexId_t *d_ind, int *const outPutTriangles,
                                      const int threads_per_block, const int number_blocks,
                                      const int shifter, const int thread_blocks, const int blockdim)
{
    (void)threads_per_block, (void)number_blocks, (void)shifter, (void)thread_blocks, (void)blockdim;
    // int32 offsets go to the library as they are (widened on the device)
    tcCheck(tc_csr_count_device32(nv, d_off, d_ind, reinterpret_cast<uint32_t *>(outPutTriangles),
                                  nullptr, nullptr));
}

// readGraphDIMACS / readGraphSNAP / readGraphMatrixMarket with the drivers'
// signatures (tests/cct.cpp:246-254): host CSR with int32 offsets, malloc'd
// (the caller frees with free()).  The last flag of each is accepted and not
// used: the result is always the symmetric, loop-free, deduplicated graph.
inline void readGraphFile(const char *path, int format, length_t **off, vertexId_t **adj,
                          length_t *nv, length_t *ne)
{
    int32_t n = 0;
    int64_t m = 0, *o = nullptr;
    int32_t *a = nullptr;
    tcCheck(tc_read_graph(path, format, 0, &n, &m, &o, &a));
    if (m > INT32_MAX) {
        tc_host_free(o);
        tc_host_free(a);
        throw cuStingerError(TC_ERR_CAPACITY, "graph has more than 2^31-1 entries; use tc_read_graph");
    }
    length_t *o32 = static_cast<length_t *>(std::malloc(sizeof(length_t) * ((size_t)n + 1)));
    if (!o32) {
        tc_host_free(o);
        tc_host_free(a);
        throw cuStingerError(TC_ERR_OOM, "host allocation failed");
    }
    for (int64_t i = 0; i <= n; ++i) o32[i] = (length_t)o[i];
    tc_host_free(o);
    *off = o32;
    *adj = a;
    *nv = n;
    *ne = (length_t)m;
}
inline void readGraphDIMACS(const char *path, length_t **off, vertexId_t **adj, length_t *nv,
                            length_t *ne, bool isRmat)
{
    (void)isRmat;
    readGraphFile(path, TC_FMT_DIMACS, off, adj, nv, ne);
}
inline void readGraphSNAP(const char *path, length_t **off, vertexId_t **adj, length_t *nv,
                          length_t *ne, bool isRmat)
{
    (void)isRmat;
    readGraphFile(path, TC_FMT_SNAP, off, adj, nv, ne);
}
inline void readGraphMatrixMarket(const char *path, length_t **off, vertexId_t **adj,
                                  length_t *nv, length_t *ne, bool undirected)
{
    (void)undirected;
    readGraphFile(path, TC_FMT_MTX, off, adj, nv, ne);
}

// ------------------------------------------------------ algorithm classes --
namespace cuStingerAlgs {

class StaticAlgorithm {
public:
    virtual ~StaticAlgorithm() {}
    virtual void Init(cuStinger &custing) = 0;
    virtual void Reset() = 0;
    virtual void Run(cuStinger &custing) = 0;
    virtual void Release() = 0;
    virtual void SyncHostWithDevice() = 0;
    virtual void SyncDeviceWithHost() = 0;
};

class StaticTriangleCounting : public StaticAlgorithm {
public:
    void Init(cuStinger &custing) override
    {
        nv_ = custing.getMaxNV();
        d_out_ = static_cast<uint32_t *>(allocDeviceArray(nv_ > 0 ? nv_ : 1, 4));
        host_.assign((size_t)nv_, 0);
        tcCheck(tc_schedule_default(&sched_));
        Reset();
    }
    void Reset() override
    {
        if (nv_ > 0) tcCheck(tc_memset(d_out_, 0, (uint64_t)nv_ * 4));
        global_ = 0;
    }
    void Run(cuStinger &custing) override
    {
        tcCheck(tc_static_count(custing.handle(), d_out_, &global_, &sched_, &stats_));
    }
    void Release() override
    {
        if (d_out_) freeDeviceArray(d_out_);
        d_out_ = nullptr;
    }
    void SyncHostWithDevice() override
    {
        if (nv_ > 0) copyArrayDeviceToHost(d_out_, host_.data(), nv_, 4);
    }
    void SyncDeviceWithHost() override
    {
        if (nv_ > 0) copyArrayHostToDevice(host_.data(), d_out_, nv_, 4);
    }
    void setSchedule(const tc_schedule &s) { sched_ = s; }
    uint64_t getGlobal() const { return global_; }
    uint64_t getTriangles() const { return global_ / 6; }
    const std::vector<uint32_t> &hostPerVertex() const { return host_; }
    const uint32_t *devicePerVertex() const { return d_out_; }
    const tc_stats &stats() const { return stats_; }

private:
    length_t nv_ = 0;
    uint32_t *d_out_ = nullptr;
    std::vector<uint32_t> host_;
    uint64_t global_ = 0;
    tc_schedule sched_;
    tc_stats stats_;
};

// Streaming triangle counting: per-vertex counts and the global count are
// kept current across batch insertions and deletions.
class StreamingTriangleCounting {
public:
    void Init(cuStinger &custing)
    {
        nv_ = custing.getMaxNV();
        d_out_ = static_cast<uint32_t *>(allocDeviceArray(nv_ > 0 ? nv_ : 1, 4));
        if (nv_ > 0) tcCheck(tc_memset(d_out_, 0, (uint64_t)nv_ * 4));
        tcCheck(tc_schedule_default(&sched_));
    }
    void runStatic(cuStinger &custing)
    {
        tcCheck(tc_static_count(custing.handle(), d_out_, &global_, &sched_, nullptr));
    }
    void batchUpdateInserted(cuStinger &custing, const int32_t *u_dev, const int32_t *v_dev,
                             int64_t B)
    {
        apply(custing, u_dev, v_dev, nullptr, B);
    }
    void batchUpdateDeleted(cuStinger &custing, const int32_t *u_dev, const int32_t *v_dev,
                            int64_t B)
    {
        std::vector<uint8_t> del((size_t)(B > 0 ? B : 1), 1);
        void *d = nullptr;
        tcCheck(tc_device_alloc(&d, (uint64_t)del.size()));
        int st = tc_memcpy(d, del.data(), del.size(), 0);
        if (st == TC_OK) {
            try {
                apply(custing, u_dev, v_dev, (const uint8_t *)d, B);
            } catch (...) {
                tc_device_free(d);
                throw;
            }
        }
        tc_device_free(d);
        tcCheck(st);
    }
    void Release()
    {
        if (d_out_) freeDeviceArray(d_out_);
        d_out_ = nullptr;
    }
    uint64_t getGlobal() const { return global_; }
    const uint32_t *devicePerVertex() const { return d_out_; }
    void copyPerVertexToHost(uint32_t *h) const { copyArrayDeviceToHost(d_out_, h, nv_, 4); }
    const tc_batch_stats &lastBatch() const { return last_; }

private:
    void apply(cuStinger &custing, const int32_t *u, const int32_t *v, const uint8_t *del, int64_t B)
    {
        int64_t dg = 0;
        tcCheck(tc_batch_apply(custing.handle(), u, v, del, B, d_out_, &dg, &sched_, &last_));
        global_ = (uint64_t)((int64_t)global_ + dg);
    }
    length_t nv_ = 0;
    uint32_t *d_out_ = nullptr;
    uint64_t global_ = 0;
    tc_schedule sched_;
    tc_batch_stats last_;
};

}  // namespace cuStingerAlgs

#endif  // CUSTINGER_TC_HPP
