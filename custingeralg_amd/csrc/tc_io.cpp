// tc_io.cpp -- graph file readers (host code): DIMACS/METIS ".graph"
// adjacency lists, SNAP edge lists and MatrixMarket coordinate files, the three
// formats the reference drivers read (tests/cct.cpp:239-254 through the
// external cuStinger utilities readGraphDIMACS / readGraphSNAP /
// readGraphMatrixMarket).  Output is a host CSR (int64 offsets, int32 ids);
// by default symmetrised, loop-free, deduplicated and row-sorted -- the form
// the triangle-counting path counts.
#include <algorithm>
#include <cctype>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/custinger_tc.h"

namespace tc {
void set_last_error(int code, const std::string &msg);
}

namespace {

// Whole file in memory, parsed with a cursor.
struct Text {
    std::vector<char> buf;
    size_t p = 0;

    bool eof() const { return p >= buf.size(); }
    void skip_line()
    {
        while (p < buf.size() && buf[p] != '\n') ++p;
        if (p < buf.size()) ++p;
    }
    // skips spaces and tabs (not newlines); returns the next char or 0
    char peek_inline()
    {
        while (p < buf.size() && (buf[p] == ' ' || buf[p] == '\t' || buf[p] == '\r')) ++p;
        return p < buf.size() ? buf[p] : 0;
    }
    // reads an unsigned integer on the current line; false at end of line/file
    bool uint_inline(int64_t &x)
    {
        const char c = peek_inline();
        if (!(c >= '0' && c <= '9')) return false;
        int64_t v = 0;
        while (p < buf.size() && buf[p] >= '0' && buf[p] <= '9') v = v * 10 + (buf[p++] - '0');
        // a fractional or exponent part (weights) is skipped
        while (p < buf.size() && (std::isalnum((unsigned char)buf[p]) || buf[p] == '.' ||
                                  buf[p] == '-' || buf[p] == '+'))
            ++p;
        x = v;
        return true;
    }
    // skips a token of any kind on the current line
    bool skip_token_inline()
    {
        const char c = peek_inline();
        if (!c || c == '\n') return false;
        while (p < buf.size() && !std::isspace((unsigned char)buf[p])) ++p;
        return true;
    }
};

int load(const char *path, Text &t)
{
    FILE *f = std::fopen(path, "rb");
    if (!f) {
        tc::set_last_error(TC_ERR_INVALID, std::string("cannot open ") + path);
        return TC_ERR_INVALID;
    }
    std::fseek(f, 0, SEEK_END);
    const long sz = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    t.buf.resize(sz > 0 ? (size_t)sz : 0);
    const size_t got = sz > 0 ? std::fread(t.buf.data(), 1, (size_t)sz, f) : 0;
    std::fclose(f);
    if ((long)got != sz) {
        tc::set_last_error(TC_ERR_INVALID, std::string("short read of ") + path);
        return TC_ERR_INVALID;
    }
    return TC_OK;
}

int bad(const std::string &what)
{
    tc::set_last_error(TC_ERR_INVALID, what);
    return TC_ERR_INVALID;
}

// DIMACS 10 / METIS: '%' comment lines, header "n m [fmt [ncon]]", then one
// line per vertex with its 1-based neighbours (vertex or edge weights, when
// fmt says so, are skipped).
int parse_dimacs(Text &t, int64_t &n, std::vector<uint64_t> &keys)
{
    while (!t.eof() && (t.peek_inline() == '%' || t.peek_inline() == '\n')) t.skip_line();
    int64_t m = 0, fmt = 0, ncon = 0;
    if (!t.uint_inline(n) || !t.uint_inline(m)) return bad("DIMACS: missing header \"n m\"");
    t.uint_inline(fmt);
    t.uint_inline(ncon);
    t.skip_line();
    const bool vw = (fmt / 10) % 10 == 1, ew = fmt % 10 == 1;
    if (vw && ncon == 0) ncon = 1;
    if (n < 0 || n > INT32_MAX) return bad("DIMACS: vertex count out of range");
    keys.reserve((size_t)std::max<int64_t>(0, 2 * m));
    for (int64_t v = 0; v < n && !t.eof();) {
        if (t.peek_inline() == '%') {
            t.skip_line();
            continue;
        }
        if (vw)
            for (int64_t c = 0; c < ncon; ++c) t.skip_token_inline();
        int64_t w;
        while (t.uint_inline(w)) {
            if (w < 1 || w > n) return bad("DIMACS: neighbour id out of range");
            keys.push_back(((uint64_t)v << 32) | (uint64_t)(w - 1));
            if (ew) t.skip_token_inline();
        }
        t.skip_line();
        ++v;
    }
    return TC_OK;
}

// SNAP: '#' comment lines, then "u v" per line with 0-based ids.
int parse_snap(Text &t, int64_t &n, std::vector<uint64_t> &keys)
{
    n = 0;
    while (!t.eof()) {
        const char c = t.peek_inline();
        if (c == '#' || c == '%' || c == '\n' || c == 0) {
            t.skip_line();
            continue;
        }
        int64_t u, v;
        if (!t.uint_inline(u) || !t.uint_inline(v)) return bad("SNAP: expected \"u v\"");
        if (u > INT32_MAX - 1 || v > INT32_MAX - 1) return bad("SNAP: id out of range");
        keys.push_back(((uint64_t)u << 32) | (uint64_t)v);
        n = std::max(n, std::max(u, v) + 1);
        t.skip_line();
    }
    return TC_OK;
}

// MatrixMarket coordinate: banner, '%' comments, "rows cols entries", then
// "i j [value]" with 1-based indices.  A symmetric banner means only one
// triangle is stored.
int parse_mtx(Text &t, int64_t &n, std::vector<uint64_t> &keys, bool &symmetric)
{
    symmetric = false;
    if (t.buf.size() >= 14 && std::strncmp(t.buf.data(), "%%MatrixMarket", 14) == 0) {
        size_t e = 0;
        while (e < t.buf.size() && t.buf[e] != '\n') ++e;
        std::string banner(t.buf.data(), e);
        for (char &c : banner) c = (char)std::tolower((unsigned char)c);
        if (banner.find("coordinate") == std::string::npos)
            return bad("MatrixMarket: only the coordinate format holds a graph");
        symmetric = banner.find("symmetric") != std::string::npos ||
                    banner.find("hermitian") != std::string::npos;
    }
    while (!t.eof() && (t.peek_inline() == '%' || t.peek_inline() == '\n')) t.skip_line();
    int64_t rows = 0, cols = 0, nz = 0;
    if (!t.uint_inline(rows) || !t.uint_inline(cols) || !t.uint_inline(nz))
        return bad("MatrixMarket: missing size line");
    t.skip_line();
    n = std::max(rows, cols);
    if (n > INT32_MAX) return bad("MatrixMarket: dimension out of range");
    keys.reserve((size_t)nz);
    while (!t.eof()) {
        const char c = t.peek_inline();
        if (c == '%' || c == '\n' || c == 0) {
            t.skip_line();
            continue;
        }
        int64_t i, j;
        if (!t.uint_inline(i) || !t.uint_inline(j)) return bad("MatrixMarket: expected \"i j\"");
        if (i < 1 || j < 1 || i > n || j > n) return bad("MatrixMarket: index out of range");
        keys.push_back(((uint64_t)(i - 1) << 32) | (uint64_t)(j - 1));
        t.skip_line();
    }
    return TC_OK;
}

}  // namespace

extern "C" {

int tc_read_graph(const char *path, int format, uint32_t flags, int32_t *nv, int64_t *nnz,
                  int64_t **off, int32_t **adj)
{
    try {
        if (!path || !nv || !nnz || !off || !adj) return bad("tc_read_graph: null argument");
        *off = nullptr;
        *adj = nullptr;
        if (format == TC_FMT_AUTO) {
            const std::string s(path);
            if (s.find(".graph") != std::string::npos)
                format = TC_FMT_DIMACS;
            else if (s.find(".txt") != std::string::npos)
                format = TC_FMT_SNAP;
            else if (s.find(".mtx") != std::string::npos)
                format = TC_FMT_MTX;
            else
                return bad("tc_read_graph: unknown graph type (expected .graph, .txt or .mtx)");
        }
        Text t;
        int st = load(path, t);
        if (st != TC_OK) return st;
        int64_t n = 0;
        std::vector<uint64_t> keys;
        bool sym_input = false;
        if (format == TC_FMT_DIMACS)
            st = parse_dimacs(t, n, keys);
        else if (format == TC_FMT_SNAP)
            st = parse_snap(t, n, keys);
        else if (format == TC_FMT_MTX)
            st = parse_mtx(t, n, keys, sym_input);
        else
            return bad("tc_read_graph: bad format");
        if (st != TC_OK) return st;
        t.buf.clear();
        t.buf.shrink_to_fit();
        const bool as_is = (flags & TC_READ_AS_IS) != 0;
        if (!as_is || sym_input) {
            // add the reverse of every pair (a symmetric MatrixMarket file stores one triangle)
            const size_t k = keys.size();
            keys.reserve(2 * k);
            for (size_t i = 0; i < k; ++i) {
                const uint64_t x = keys[i];
                keys.push_back((x << 32) | (x >> 32));
            }
        }
        if (!as_is)
            keys.erase(std::remove_if(keys.begin(), keys.end(),
                                      [](uint64_t x) { return (x >> 32) == (x & 0xffffffffu); }),
                       keys.end());
        std::sort(keys.begin(), keys.end());
        keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
        int64_t *o = static_cast<int64_t *>(std::malloc(sizeof(int64_t) * (size_t)(n + 1)));
        int32_t *a = static_cast<int32_t *>(std::malloc(sizeof(int32_t) * std::max<size_t>(keys.size(), 1)));
        if (!o || !a) {
            std::free(o);
            std::free(a);
            tc::set_last_error(TC_ERR_OOM, "tc_read_graph: host allocation failed");
            return TC_ERR_OOM;
        }
        std::fill(o, o + n + 1, 0);
        for (size_t i = 0; i < keys.size(); ++i) {
            o[(keys[i] >> 32) + 1]++;
            a[i] = (int32_t)(keys[i] & 0xffffffffu);
        }
        for (int64_t v = 0; v < n; ++v) o[v + 1] += o[v];
        *nv = (int32_t)n;
        *nnz = (int64_t)keys.size();
        *off = o;
        *adj = a;
        return TC_OK;
    } catch (const std::bad_alloc &) {
        tc::set_last_error(TC_ERR_OOM, "tc_read_graph: host allocation failed");
        return TC_ERR_OOM;
    }
}

int tc_host_free(void *p)
{
    std::free(p);
    return TC_OK;
}

}  // extern "C"
