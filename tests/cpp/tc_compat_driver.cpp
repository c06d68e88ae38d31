// tc_compat_driver.cpp -- drives the C++ host layer (include/custinger_tc.hpp)
// the way the reference's drivers drive cuStingerAlg: the launch-parameter
// sweep of tests/cct.cpp:177-209 over callDeviceAllTriangles and
// callDeviceAllTrianglesCSR, and the InsertionTest flow of
// tests/cct_upd.cpp:194-414 (remove edges, re-insert them in batches with
// planted duplicates, count new triangles with callDeviceNewTriangles).
// Every result is checked against the CPU oracle; exit status 0 = all equal.
#include <cstdio>
#include <cstdlib>
#include <set>
#include <utility>
#include <vector>

#include "../../include/custinger_tc.hpp"
#include "../../oracle/tc_oracle.h"

static int g_fail = 0;
#define CHECK(cond, ...)                                                                           \
    do {                                                                                           \
        if (!(cond)) {                                                                             \
            std::printf("FAIL %s:%d ", __FILE__, __LINE__);                                        \
            std::printf(__VA_ARGS__);                                                              \
            std::printf("\n");                                                                     \
            ++g_fail;                                                                              \
        }                                                                                          \
    } while (0)

struct Csr {
    int32_t nv = 0;
    std::vector<int64_t> off;
    std::vector<int32_t> adj;
};

static Csr rmat(int scale, int ef, uint64_t seed)
{
    Csr c;
    c.nv = 1 << scale;
    const int64_t m = (int64_t)ef << scale;
    std::vector<int32_t> u((size_t)m), v((size_t)m);
    uint64_t t[3];
    tco_rmat_thresholds(0.57, 0.19, 0.19, t);
    tco_gen_rmat_pairs(scale, m, t, seed, u.data(), v.data());
    c.off.assign((size_t)c.nv + 1, 0);
    c.adj.assign((size_t)(2 * m), 0);
    const int64_t nnz = tco_canon_csr(c.nv, m, u.data(), v.data(), c.off.data(), c.adj.data());
    c.adj.resize((size_t)nnz);
    return c;
}

static Csr from_set(int32_t nv, const std::set<std::pair<int32_t, int32_t>> &e)
{
    Csr c;
    c.nv = nv;
    c.off.assign((size_t)nv + 1, 0);
    for (auto &p : e) c.off[(size_t)p.first + 1]++;
    for (int32_t i = 0; i < nv; ++i) c.off[(size_t)i + 1] += c.off[(size_t)i];
    for (auto &p : e) c.adj.push_back(p.second);  // std::set iterates in (src,dst) order
    return c;
}

static void init_custinger(cuStinger &cs, const Csr &c, std::vector<length_t> &off32)
{
    off32.assign(c.off.begin(), c.off.end());
    cuStingerInitConfig cuInit;
    cuInit.initState = eInitStateCSR;
    cuInit.maxNV = c.nv + 1;
    cuInit.useVWeight = false;
    cuInit.isSemantic = false;
    cuInit.useEWeight = false;
    cuInit.csrNV = c.nv;
    cuInit.csrNE = (length_t)c.adj.size();
    cuInit.csrOff = off32.data();
    cuInit.csrAdj = c.adj.data();
    cs.initializeCuStinger(cuInit);
}

static std::vector<uint32_t> oracle_pv(const Csr &c)
{
    std::vector<uint32_t> o((size_t)std::max(c.nv, 1));
    tco_custinger_count(c.nv, c.off.data(), c.adj.data(), o.data(), 4);
    o.resize((size_t)c.nv);
    return o;
}

static void static_sweep()
{
    const Csr c = rmat(12, 16, 2);
    const int32_t nv = c.nv;
    std::vector<length_t> off32;
    cuStinger custing(defaultInitAllocater, defaultUpdateAllocater);
    init_custinger(custing, c, off32);
    CHECK(custing.getNumberEdgesUsed() == (int64_t)c.adj.size(), "edges used");
    const std::vector<uint32_t> want = oracle_pv(c);
    std::vector<uint32_t> want_csr((size_t)nv);
    tco_csr_count(nv, c.off.data(), c.adj.data(), want_csr.data(), 4);

    triangle_t *d_triangles = (triangle_t *)allocDeviceArray(nv + 1, sizeof(triangle_t));
    length_t *d_off = (length_t *)allocDeviceArray(nv + 1, sizeof(length_t));
    vertexId_t *d_ind = (vertexId_t *)allocDeviceArray((length_t)c.adj.size(), sizeof(vertexId_t));
    copyArrayHostToDevice(off32.data(), d_off, nv + 1, sizeof(length_t));
    copyArrayHostToDevice(c.adj.data(), d_ind, (length_t)c.adj.size(), sizeof(vertexId_t));
    std::vector<triangle_t> h((size_t)nv + 1);
    const int blockSizes[] = {32, 64, 96, 128, 192, 256};
    const int tpis[] = {1, 2, 4, 8, 16, 32};
    const int shifts[] = {0, 1, 2, 3, 4, 5};
    int points = 0;
    for (int sps : blockSizes)
        for (int t = 0; t < 6; ++t) {
            const int tsp = tpis[t], shifter = shifts[t], nbl = sps / tsp, blocks = 16000;
            callDeviceAllTrianglesCSR(nv, d_off, d_ind, d_triangles, tsp, nbl, shifter, blocks, sps);
            copyArrayDeviceToHost(d_triangles, h.data(), nv, sizeof(triangle_t));
            for (int32_t i = 0; i < nv; ++i)
                if ((uint32_t)h[(size_t)i] != want_csr[(size_t)i]) {
                    CHECK(false, "CSR variant vertex %d at tpi %d", i, tsp);
                    break;
                }
            callDeviceAllTriangles(custing, d_triangles, tsp, nbl, shifter, blocks, sps);
            copyArrayDeviceToHost(d_triangles, h.data(), nv, sizeof(triangle_t));
            for (int32_t i = 0; i < nv; ++i)
                if ((uint32_t)h[(size_t)i] != want[(size_t)i]) {
                    CHECK(false, "cuStinger variant vertex %d at tpi %d", i, tsp);
                    break;
                }
            // sums as the driver prints them: 6T and 3T
            ++points;
        }
    const int64_t s6 = tco_sum_triangle_array(h.data(), nv);
    const uint64_t T = tco_forward_count(nv, c.off.data(), c.adj.data(), nullptr, 4);
    CHECK(s6 == (int64_t)(6 * T), "sum %lld vs 6T %llu", (long long)s6, (unsigned long long)(6 * T));
    std::printf("static sweep: %d launch configurations, nv=%d nnz=%zu, 6T=%lld\n", points, nv,
                c.adj.size(), (long long)s6);

    cuStingerAlgs::StaticTriangleCounting alg;
    alg.Init(custing);
    alg.Run(custing);
    alg.SyncHostWithDevice();
    CHECK(alg.hostPerVertex() == want, "StaticTriangleCounting per-vertex");
    CHECK(alg.getGlobal() == 6 * T, "StaticTriangleCounting global");
    alg.Reset();
    alg.Release();
    freeDeviceArray(d_triangles);
    freeDeviceArray(d_off);
    freeDeviceArray(d_ind);
    custing.freecuStinger();
}

static void insertion_test()
{
    const Csr orig = rmat(11, 16, 3);
    const int32_t nv = orig.nv;
    std::set<std::pair<int32_t, int32_t>> edges;
    for (int32_t a = 0; a < nv; ++a)
        for (int64_t k = orig.off[(size_t)a]; k < orig.off[(size_t)a + 1]; ++k)
            edges.insert({a, orig.adj[(size_t)k]});
    // remove numEdges random edges per batch, plant in-batch repeats and edges
    // that stay in the graph (tests/cct_upd.cpp:274-342)
    const int numBatches = 5, numEdges = 400, dupEdges = 40, dupsFromCUS = 40;
    std::vector<std::vector<std::pair<int32_t, int32_t>>> batches((size_t)numBatches);
    std::vector<std::pair<int32_t, int32_t>> all(edges.begin(), edges.end());
    uint64_t ctr = 0;
    auto rnd = [&](uint64_t n) { return tco_rng(100, ctr++) % n; };
    std::set<std::pair<int32_t, int32_t>> removed;
    for (int b = 0; b < numBatches; ++b) {
        auto &B = batches[(size_t)b];
        for (int i = 0; i < numEdges; ++i) {
            std::pair<int32_t, int32_t> e;
            do {
                e = all[(size_t)rnd(all.size())];
            } while (removed.count(e) || removed.count({e.second, e.first}));
            removed.insert(e);
            B.push_back(e);
        }
        for (int i = 0; i < dupEdges; ++i) B.push_back(B[(size_t)i]);
        for (int i = 0; i < dupsFromCUS; ++i) {
            std::pair<int32_t, int32_t> e;
            do {
                e = all[(size_t)rnd(all.size())];
            } while (removed.count(e) || removed.count({e.second, e.first}));
            B.push_back(e);
        }
    }
    std::set<std::pair<int32_t, int32_t>> reduced = edges;
    for (auto &e : removed) {
        reduced.erase(e);
        reduced.erase({e.second, e.first});
    }
    const Csr red = from_set(nv, reduced);
    const std::vector<uint32_t> pv_orig = oracle_pv(orig), pv_red = oracle_pv(red);
    const uint64_t T_orig = tco_forward_count(nv, orig.off.data(), orig.adj.data(), nullptr, 4);
    const uint64_t T_red = tco_forward_count(nv, red.off.data(), red.adj.data(), nullptr, 4);

    std::vector<length_t> off32;
    cuStinger custingTest(defaultInitAllocater, defaultUpdateAllocater);
    init_custinger(custingTest, red, off32);
    triangle_t *d_new = (triangle_t *)allocDeviceArray(nv + 1, sizeof(triangle_t));
    std::vector<triangle_t> zero((size_t)nv + 1, 0);
    copyArrayHostToDevice(zero.data(), d_new, nv + 1, sizeof(triangle_t));
    length_t total_allocs = 0;
    std::set<std::pair<int32_t, int32_t>> store = reduced;  // the store's entries before each batch
    for (int b = 0; b < numBatches; ++b) {
        const auto &B = batches[(size_t)b];
        const length_t nt = (length_t)B.size();
        BatchUpdateData bud(nt * 2, true, nv);
        for (length_t i = 0; i < nt; ++i) {
            bud.getSrc()[i] = B[(size_t)i].first;
            bud.getDst()[i] = B[(size_t)i].second;
            bud.getSrc()[nt + i] = B[(size_t)i].second;  // reversed copies (:337-341)
            bud.getDst()[nt + i] = B[(size_t)i].first;
        }
        // expected getvNumDuplicates: per source, the entries the insertion
        // leaves alone (repeats after the first, entries already present)
        std::vector<length_t> want_dup((size_t)nv, 0);
        {
            std::set<std::pair<int32_t, int32_t>> seen;
            for (length_t i = 0; i < 2 * nt; ++i) {
                const std::pair<int32_t, int32_t> e{bud.getSrc()[i], bud.getDst()[i]};
                if (store.count(e) || !seen.insert(e).second) want_dup[(size_t)e.first]++;
            }
            store.insert(seen.begin(), seen.end());
        }
        for (int32_t i = 0; i < nv; ++i) bud.getvNumDuplicates()[i] = 0;  // tests/cct_upd.cpp:335
        BatchUpdate bu1(bud);
        length_t allocs = 0;
        bu1.sortDeviceBUD(128);
        custingTest.edgeInsertionsSorted(bu1, allocs);
        bool dup_ok = true;
        for (int32_t i = 0; i < nv; ++i) dup_ok = dup_ok && bud.getvNumDuplicates()[i] == want_dup[(size_t)i];
        CHECK(dup_ok, "getvNumDuplicates after batch %d", b);
        callDeviceNewTriangles(custingTest, bu1, d_new, 1, 128, 0, 16000, 128, nullptr, nullptr);
        total_allocs += allocs;
    }
    std::vector<triangle_t> h((size_t)nv);
    copyArrayDeviceToHost(d_new, h.data(), nv, sizeof(triangle_t));
    const int64_t sum_new = tco_sum_triangle_array(h.data(), nv);
    CHECK(sum_new == (int64_t)(6 * T_orig) - (int64_t)(6 * T_red), "Triangles created by batch update");
    for (int32_t i = 0; i < nv; ++i)
        if (pv_red[(size_t)i] + (uint32_t)h[(size_t)i] != pv_orig[(size_t)i]) {
            CHECK(false, "per-vertex delta at %d", i);
            break;
        }
    CHECK(custingTest.getNumberEdgesUsed() == (int64_t)orig.adj.size(), "edges after re-insertion");
    std::printf("insertion test: should be %lld, got %lld (allocs %d, used/allocated %lld/%lld)\n",
                (long long)(6 * T_orig - 6 * T_red), (long long)sum_new, total_allocs,
                (long long)custingTest.getNumberEdgesUsed(),
                (long long)custingTest.getNumberEdgesAllocated());
    freeDeviceArray(d_new);

    // the streaming class over the same graph: delete the batches again
    cuStingerAlgs::StreamingTriangleCounting st;
    st.Init(custingTest);
    st.runStatic(custingTest);
    CHECK(st.getGlobal() == 6 * T_orig, "runStatic global");
    for (int b = 0; b < numBatches; ++b) {
        std::vector<int32_t> u, v;
        for (int i = 0; i < numEdges; ++i) {
            u.push_back(batches[(size_t)b][(size_t)i].first);
            v.push_back(batches[(size_t)b][(size_t)i].second);
        }
        int32_t *du = (int32_t *)allocDeviceArray((length_t)u.size(), 4);
        int32_t *dv = (int32_t *)allocDeviceArray((length_t)v.size(), 4);
        copyArrayHostToDevice(u.data(), du, (length_t)u.size(), 4);
        copyArrayHostToDevice(v.data(), dv, (length_t)v.size(), 4);
        st.batchUpdateDeleted(custingTest, du, dv, (int64_t)u.size());
        freeDeviceArray(du);
        freeDeviceArray(dv);
    }
    CHECK(st.getGlobal() == 6 * T_red, "streaming global after deletions");
    std::vector<uint32_t> hp((size_t)nv);
    st.copyPerVertexToHost(hp.data());
    CHECK(hp == pv_red, "streaming per-vertex after deletions");
    st.Release();
    custingTest.freecuStinger();
}

// tests/katz-tester.cu:211-252 on a directed graph: a directed batch inserted
// with edgeInsertions and removed with edgeDeletions; after each step the
// store equals the expected one (compareCUS against a store built from the
// expected CSR), and the static count of the directed store equals the
// reference formula's.
static void directed_test()
{
    const int32_t nv = 2000;
    std::set<std::pair<int32_t, int32_t>> edges;
    uint64_t ctr = 0;
    auto rnd = [&](uint64_t n) { return (int32_t)(tco_rng(200, ctr++) % n); };
    while (edges.size() < 20000) {
        const int32_t a = rnd(nv), b = rnd(nv);
        if (a != b) edges.insert({a, b});  // one direction only: a directed graph
    }
    const Csr g0 = from_set(nv, edges);
    std::vector<length_t> off32;
    cuStinger custing(defaultInitAllocater, defaultUpdateAllocater);
    init_custinger(custing, g0, off32);
    CHECK(custing.info().symmetric == 0, "directed store flagged asymmetric");
    // generateEdgeUpdates shape: uniform directed entries, repeats allowed
    const length_t nb = 3000;
    BatchUpdateData bud(nb, true);
    std::set<std::pair<int32_t, int32_t>> batch;
    for (length_t i = 0; i < nb; ++i) {
        int32_t a = rnd(nv), b = rnd(nv);
        if (a == b) b = (b + 1) % nv;
        if (i % 10 == 0 && i > 0) {  // every tenth entry repeats an earlier one
            a = bud.getSrc()[i / 2];
            b = bud.getDst()[i / 2];
        }
        bud.getSrc()[i] = a;
        bud.getDst()[i] = b;
        batch.insert({a, b});
    }
    BatchUpdate bu(bud);
    length_t allocs = 0;
    custing.edgeInsertions(bu, allocs);
    std::set<std::pair<int32_t, int32_t>> after = edges;
    after.insert(batch.begin(), batch.end());
    const Csr g1 = from_set(nv, after);
    {
        cuStinger want;
        std::vector<length_t> o;
        init_custinger(want, g1, o);
        CHECK(compareCUS(&custing, &want) == 0, "directed insertions");
        CHECK(compareCUS(&custing, &custing) == 0, "a store equals itself");
    }
    CHECK(custing.getNumberEdgesUsed() == (int64_t)after.size(), "entries after directed insertions");
    // counting a directed store: the reference formula over the rows as they are
    triangle_t *d = (triangle_t *)allocDeviceArray(nv, sizeof(triangle_t));
    callDeviceAllTriangles(custing, d, 1, 128, 0, 16000, 128);
    std::vector<triangle_t> h((size_t)nv);
    copyArrayDeviceToHost(d, h.data(), nv, sizeof(triangle_t));
    const std::vector<uint32_t> want_pv = oracle_pv(g1);
    bool same = true;
    for (int32_t i = 0; i < nv; ++i) same = same && (uint32_t)h[(size_t)i] == want_pv[(size_t)i];
    CHECK(same, "count of the directed store");
    freeDeviceArray(d);
    custing.edgeDeletions(bu);
    std::set<std::pair<int32_t, int32_t>> left = edges;
    for (auto &e : batch) left.erase(e);
    const Csr g2 = from_set(nv, left);
    {
        cuStinger want;
        std::vector<length_t> o;
        init_custinger(want, g2, o);
        CHECK(compareCUS(&custing, &want) == 0, "directed deletions");
        cuStinger other;
        init_custinger(other, g0, o);
        const bool differs = !batch.empty() && left.size() != edges.size();
        CHECK(!differs || compareCUS(&custing, &other) > 0, "compareCUS finds a difference");
    }
    std::printf("directed test: %zu entries, batch %d (%zu distinct), %zu after insert, %zu after "
                "delete, allocs %d\n",
                edges.size(), nb, batch.size(), after.size(), left.size(), allocs);
    custing.freecuStinger();
}

// readGraphDIMACS -> initializeCuStinger -> callDeviceAllTriangles, as the
// drivers do with a file argument (tests/cct.cpp:246-277).
static void file_test()
{
    const Csr c = rmat(9, 8, 5);
    char path[] = "/tmp/tc_compat_XXXXXX.graph";
    const int fd = mkstemps(path, 6);
    CHECK(fd >= 0, "temporary file");
    if (fd < 0) return;
    FILE *f = fdopen(fd, "w");
    std::fprintf(f, "%% generated\n%d %lld\n", c.nv, (long long)(c.adj.size() / 2));
    for (int32_t v = 0; v < c.nv; ++v) {
        for (int64_t k = c.off[(size_t)v]; k < c.off[(size_t)v + 1]; ++k)
            std::fprintf(f, "%d ", c.adj[(size_t)k] + 1);
        std::fprintf(f, "\n");
    }
    std::fclose(f);
    length_t nv = 0, ne = 0, *off = nullptr;
    vertexId_t *adj = nullptr;
    readGraphDIMACS(path, &off, &adj, &nv, &ne, false);
    std::remove(path);
    CHECK(nv == c.nv && ne == (length_t)c.adj.size(), "file sizes");
    cuStingerInitConfig cuInit;
    cuInit.initState = eInitStateCSR;
    cuInit.maxNV = nv + 1;
    cuInit.csrNV = nv;
    cuInit.csrNE = ne;
    cuInit.csrOff = off;
    cuInit.csrAdj = adj;
    cuStinger custing(defaultInitAllocater, defaultUpdateAllocater);
    custing.initializeCuStinger(cuInit);
    triangle_t *d = (triangle_t *)allocDeviceArray(nv + 1, sizeof(triangle_t));
    callDeviceAllTriangles(custing, d, 1, 128, 0, 16000, 128);
    std::vector<triangle_t> h((size_t)nv);
    copyArrayDeviceToHost(d, h.data(), nv, sizeof(triangle_t));
    const std::vector<uint32_t> want = oracle_pv(c);
    bool same = true;
    for (int32_t i = 0; i < nv; ++i) same = same && (uint32_t)h[(size_t)i] == want[(size_t)i];
    CHECK(same, "count of the graph read from a DIMACS file");
    std::printf("file test: nv=%d ne=%d sum=%lld\n", nv, ne, (long long)tco_sum_triangle_array(h.data(), nv));
    freeDeviceArray(d);
    std::free(off);
    std::free(adj);
    custing.freecuStinger();
}

int main()
{
    try {
        static_sweep();
        insertion_test();
        directed_test();
        file_test();
    } catch (const cuStingerError &e) {
        std::printf("FAIL exception: %s\n", e.what());
        return 2;
    }
    if (g_fail) {
        std::printf("FAILED checks: %d\n", g_fail);
        return 1;
    }
    std::printf("compat driver ok\n");
    return 0;
}
