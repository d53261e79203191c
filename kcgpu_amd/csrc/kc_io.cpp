// kc_io.cpp -- edge-list loaders and the host CSR builder.
//
// Formats follow graph::EdgeListFile (include/FIleReader.cuh:33-476):
//   .mtx       Matrix Market coordinate; '%' lines skipped, flags read from
//              '%%' lines (:321-375), entries "row col [value]" (:149-234)
//   .tsv/.txt  "dst src" per line, '%' and '#' lines skipped (:101-146)
//   .bel       24-byte records: dst u64 @0, src u64 @8, weight u64 @16 (:51-95)
// kc_load builds the simple undirected graph from whatever the file holds;
// read_edges_reference reproduces the reference's raw read loop, defects
// included, for documentation tests.
#include <algorithm>
#include <atomic>
#include <cerrno>
#include <cstdarg>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include <omp.h>

#include "kc_internal.h"

namespace kc {

static thread_local std::string g_err;

void set_error(const std::string &msg) { g_err = msg; }

int fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

const char *last_error() { return g_err.c_str(); }

// ---------------------------------------------------------------------------
// CSR builder.  CSRCOO::from_edgelist (include/CSRCOO.cuh:37-80) needs a
// src-sorted list, drops self-loops and keeps duplicates; this builder sorts,
// symmetrizes, drops self-loops and duplicates, so the result is the CSR of
// the simple undirected graph.  Parallel: degree count, scan, scatter, then a
// per-row sort + unique + compaction.
// ---------------------------------------------------------------------------
int build_csr(uint32_t n, size_t ne, const uint32_t *src, const uint32_t *dst, HostCSR &out)
{
    uint64_t maxid = 0;
    bool any = false;
#pragma omp parallel for reduction(max : maxid) reduction(|| : any)
    for (size_t i = 0; i < ne; ++i) {
        if (src[i] == dst[i]) continue;
        any = true;
        uint64_t a = src[i] > dst[i] ? src[i] : dst[i];
        if (a > maxid) maxid = a;
    }
    uint64_t nn = n;
    if (nn == 0) nn = any ? maxid + 1 : 0;
    if (any && maxid >= nn) return fail(KC_ERR_ARG, "vertex id %llu >= n=%llu", (unsigned long long)maxid, (unsigned long long)nn);
    if (nn >= 0xFFFFFFFFull) return fail(KC_ERR_LIMIT, "n=%llu does not fit 32-bit ids", (unsigned long long)nn);

    std::vector<uint32_t> deg(nn + 1, 0);
#pragma omp parallel for
    for (size_t i = 0; i < ne; ++i) {
        uint32_t a = src[i], b = dst[i];
        if (a == b) continue;
        __atomic_fetch_add(&deg[a], 1u, __ATOMIC_RELAXED);
        __atomic_fetch_add(&deg[b], 1u, __ATOMIC_RELAXED);
    }
    std::vector<uint64_t> off(nn + 1, 0);
    for (uint64_t v = 0; v < nn; ++v) off[v + 1] = off[v] + deg[v];
    uint64_t tot = off[nn];
    std::vector<uint32_t> tmp;
    try { tmp.resize(tot); } catch (...) { return fail(KC_ERR_NOMEM, "CSR build: %llu entries", (unsigned long long)tot); }
    std::vector<uint64_t> pos(off.begin(), off.end() - 1);
#pragma omp parallel for
    for (size_t i = 0; i < ne; ++i) {
        uint32_t a = src[i], b = dst[i];
        if (a == b) continue;
        tmp[__atomic_fetch_add(&pos[a], 1ull, __ATOMIC_RELAXED)] = b;
        tmp[__atomic_fetch_add(&pos[b], 1ull, __ATOMIC_RELAXED)] = a;
    }
    // sort + unique each row; deg[] becomes the unique count
#pragma omp parallel for schedule(dynamic, 256)
    for (uint64_t v = 0; v < nn; ++v) {
        uint32_t *b = tmp.data() + off[v], *e = tmp.data() + off[v + 1];
        std::sort(b, e);
        deg[v] = (uint32_t)(std::unique(b, e) - b);
    }
    out.n = (uint32_t)nn;
    out.rowptr.assign(nn + 1, 0);
    uint64_t s = 0;
    for (uint64_t v = 0; v < nn; ++v) {
        out.rowptr[v] = (uint32_t)s;
        s += deg[v];
    }
    if (s >= 0xFFFFFFFFull) return fail(KC_ERR_LIMIT, "%llu CSR entries exceed 32-bit offsets", (unsigned long long)s);
    out.rowptr[nn] = (uint32_t)s;
    out.col.resize(s);
#pragma omp parallel for schedule(dynamic, 1024)
    for (uint64_t v = 0; v < nn; ++v)
        std::memcpy(out.col.data() + out.rowptr[v], tmp.data() + off[v], sizeof(uint32_t) * deg[v]);
    return KC_OK;
}

uint64_t fnv1a_csr(const HostCSR &g)
{
    uint64_t h = 0xcbf29ce484222325ull;
    auto mix = [&h](uint32_t x) {
        for (int i = 0; i < 4; ++i) {
            h ^= (x >> (8 * i)) & 0xFF;
            h *= 0x100000001b3ull;
        }
    };
    for (uint32_t x : g.rowptr) mix(x);
    for (uint32_t x : g.col) mix(x);
    return h;
}

// ---------------------------------------------------------------------------
// file helpers
// ---------------------------------------------------------------------------
static bool ends_with(const std::string &s, const char *suf)
{
    size_t l = std::strlen(suf);
    return s.size() >= l && s.compare(s.size() - l, l, suf) == 0;
}

static int read_all(const char *path, std::string &buf)
{
    FILE *f = std::fopen(path, "rb");
    if (!f) return fail(KC_ERR_IO, "cannot open '%s': %s", path, std::strerror(errno));
    std::fseek(f, 0, SEEK_END);
    long len = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    if (len < 0) { std::fclose(f); return fail(KC_ERR_IO, "cannot size '%s'", path); }
    buf.resize((size_t)len);
    size_t got = len ? std::fread(&buf[0], 1, (size_t)len, f) : 0;
    std::fclose(f);
    if (got != (size_t)len) return fail(KC_ERR_IO, "short read on '%s'", path);
    return KC_OK;
}

// Parse an unsigned decimal at p (after optional blanks); returns false if none.
static bool parse_u64(const char *&p, const char *end, uint64_t &v, bool allow_neg_error, bool &neg)
{
    while (p < end && (*p == ' ' || *p == '\t' || *p == '\r')) ++p;
    neg = false;
    if (p < end && (*p == '-' || *p == '+')) {
        neg = (*p == '-');
        ++p;
    }
    (void)allow_neg_error;
    if (p >= end || *p < '0' || *p > '9') return false;
    uint64_t x = 0;
    while (p < end && *p >= '0' && *p <= '9') {
        x = x * 10 + (uint64_t)(*p - '0');
        ++p;
    }
    v = x;
    return true;
}

struct EdgeBuf {
    std::vector<uint32_t> s, d;
};

static int add_edge(EdgeBuf &eb, uint64_t a, uint64_t b, const char *path, size_t line)
{
    if (a >= 0xFFFFFFFFull || b >= 0xFFFFFFFFull)
        return fail(KC_ERR_FORMAT, "%s:%zu: vertex id does not fit 32 bits", path, line);
    if (a == b) return KC_OK;
    eb.s.push_back((uint32_t)a);
    eb.d.push_back((uint32_t)b);
    return KC_OK;
}

// Matrix Market (include/FIleReader.cuh:149-234, :321-375).
static int load_mtx(const char *path, EdgeBuf &eb)
{
    std::string buf;
    int rc = read_all(path, buf);
    if (rc) return rc;
    const char *p = buf.data(), *end = p + buf.size();
    bool header_done = false, array = false;
    size_t line = 0;
    while (p < end) {
        const char *e = (const char *)std::memchr(p, '\n', (size_t)(end - p));
        if (!e) e = end;
        ++line;
        const char *q = p;
        p = e + 1;
        while (q < e && (*q == ' ' || *q == '\t' || *q == '\r')) ++q;
        if (q == e) continue;  // blank line
        if (*q == '%') {
            if (!header_done && e - q >= 2 && q[1] == '%') {
                std::string l(q, e);
                array = l.find("array") != std::string::npos;
            }
            continue;
        }
        if (!header_done) {
            uint64_t r, c, nz;
            bool neg;
            if (!parse_u64(q, e, r, true, neg) || !parse_u64(q, e, c, true, neg))
                return fail(KC_ERR_FORMAT, "%s:%zu: bad Matrix Market size line", path, line);
            if (array) return fail(KC_ERR_FORMAT, "%s: dense 'array' Matrix Market is not a graph", path);
            if (!parse_u64(q, e, nz, true, neg))
                return fail(KC_ERR_FORMAT, "%s:%zu: size line needs rows cols entries", path, line);
            if (r != c) return fail(KC_ERR_FORMAT, "%s: matrix is not square (%llu x %llu)", path,
                                    (unsigned long long)r, (unsigned long long)c);
            header_done = true;
            eb.s.reserve(nz);
            eb.d.reserve(nz);
            continue;
        }
        uint64_t a, b;
        bool na, nb;
        if (!parse_u64(q, e, a, true, na) || !parse_u64(q, e, b, true, nb) || na || nb)
            return fail(KC_ERR_FORMAT, "%s:%zu: bad entry", path, line);
        if ((rc = add_edge(eb, a, b, path, line))) return rc;
    }
    if (!header_done) return fail(KC_ERR_FORMAT, "%s: no Matrix Market size line", path);
    return KC_OK;
}

// TSV/TXT "dst src" (include/FIleReader.cuh:101-146).
static int load_tsv(const char *path, EdgeBuf &eb)
{
    std::string buf;
    int rc = read_all(path, buf);
    if (rc) return rc;
    const char *p = buf.data(), *end = p + buf.size();
    size_t line = 0;
    while (p < end) {
        const char *e = (const char *)std::memchr(p, '\n', (size_t)(end - p));
        if (!e) e = end;
        ++line;
        const char *q = p;
        p = e + 1;
        if (q < e && (*q == '%' || *q == '#')) continue;
        const char *t = q;
        while (t < e && (*t == ' ' || *t == '\t' || *t == '\r')) ++t;
        if (t == e) continue;  // blank line
        uint64_t d, s;
        bool nd, ns;
        if (!parse_u64(q, e, d, true, nd) || !parse_u64(q, e, s, true, ns) || nd || ns)
            return fail(KC_ERR_FORMAT, "%s:%zu: expected two vertex ids", path, line);
        if ((rc = add_edge(eb, s, d, path, line))) return rc;
    }
    return KC_OK;
}

// BEL 24-byte records (include/FIleReader.cuh:51-95).
static int load_bel(const char *path, EdgeBuf &eb)
{
    std::string buf;
    int rc = read_all(path, buf);
    if (rc) return rc;
    if (buf.size() % 24 != 0) return fail(KC_ERR_FORMAT, "%s: size %zu is not a multiple of 24", path, buf.size());
    size_t nrec = buf.size() / 24;
    eb.s.reserve(nrec);
    eb.d.reserve(nrec);
    for (size_t i = 0; i < nrec; ++i) {
        uint64_t dst, src;
        std::memcpy(&dst, buf.data() + 24 * i + 0, 8);
        std::memcpy(&src, buf.data() + 24 * i + 8, 8);
        if ((rc = add_edge(eb, src, dst, path, i + 1))) return rc;
    }
    return KC_OK;
}

int load_file(const char *path, HostCSR &out)
{
    if (!path) return fail(KC_ERR_ARG, "null path");
    std::string s(path);
    EdgeBuf eb;
    int rc;
    if (ends_with(s, ".bel")) rc = load_bel(path, eb);
    else if (ends_with(s, ".tsv") || ends_with(s, ".txt")) rc = load_tsv(path, eb);
    else if (ends_with(s, ".mtx")) rc = load_mtx(path, eb);
    else return fail(KC_ERR_FORMAT, "no reader for file '%s' (expected .mtx, .tsv, .txt or .bel)", path);
    if (rc) return rc;
    return build_csr(0, eb.s.size(), eb.s.data(), eb.d.data(), out);
}

// ---------------------------------------------------------------------------
// The reference's raw read loop (src/main.cu:84-87): get_edges(v, 100)
// resizes v to 100 (keeping old contents), the reader fills it, v is resized
// to the number read, and the loop stops at the first batch that reads 0.
// ---------------------------------------------------------------------------
namespace {
struct RefReader {
    std::string buf;
    size_t p = 0;
    bool getline(std::string &l)  // std::getline: false at EOF with nothing read
    {
        if (p >= buf.size()) { l.clear(); return false; }
        size_t e = buf.find('\n', p);
        if (e == std::string::npos) e = buf.size();
        l.assign(buf, p, e - p);
        p = e + 1;
        return true;
    }
};
}  // namespace

int read_edges_reference(const char *path, std::vector<uint32_t> &src, std::vector<uint32_t> &dst)
{
    if (!path) return fail(KC_ERR_ARG, "null path");
    std::string s(path);
    RefReader rr;
    int rc = read_all(path, rr.buf);
    if (rc) return rc;
    enum { BEL, TSV, MTX } type;
    if (ends_with(s, ".bel")) type = BEL;
    else if (ends_with(s, ".tsv") || ends_with(s, ".txt")) type = TSV;
    else if (ends_with(s, ".mtx")) type = MTX;
    else return fail(KC_ERR_FORMAT, "no reader for file '%s'", path);

    bool symmetric = false, array = false, pattern = false;
    if (type == MTX) {  // AdvanceWithMarketFile (:321-375): flags from every '%%' line
        std::string l;
        while (true) {
            rr.getline(l);
            if (l.empty() || l[0] != '%') break;
            if (l.size() >= 2 && l[1] == '%') {
                symmetric = l.find("symmetric") != std::string::npos;
                array = l.find("array") != std::string::npos;
                pattern = l.find("pattern") != std::string::npos;
            }
        }
    }
    std::vector<std::pair<uint32_t, uint32_t>> batch;  // fileEdges
    size_t belpos = 0;
    std::string tsv_line;  // `char line[100]` keeps its content between iterations
    for (;;) {
        batch.resize(100);
        size_t i = 0;
        if (type == BEL) {
            size_t avail = (rr.buf.size() - belpos) / 24;
            size_t nr = avail < 100 ? avail : 100;
            for (; i < nr; ++i) {
                uint64_t d, sr;
                std::memcpy(&d, rr.buf.data() + belpos + 24 * i, 8);
                std::memcpy(&sr, rr.buf.data() + belpos + 24 * i + 8, 8);
                batch[i] = {(uint32_t)sr, (uint32_t)d};
            }
            belpos += 24 * nr;
        } else if (type == TSV) {
            for (int j = 0; j < 50; ++j) {
                if (rr.p >= rr.buf.size()) continue;  // fscanf returns EOF
                size_t e = rr.buf.find('\n', rr.p);
                if (e == std::string::npos) e = rr.buf.size();
                if (e == rr.p)  // "%[^\n]" matches nothing: the reference re-reads the stale line forever
                    return fail(KC_ERR_FORMAT, "%s: blank line at byte %zu; the reference read loop never terminates here",
                                path, rr.p);
                tsv_line.assign(rr.buf, rr.p, e - rr.p);
                rr.p = e + 1;
                if (tsv_line[0] == '%' || tsv_line[0] == '#') continue;
                unsigned long long d, sr;
                if (std::sscanf(tsv_line.c_str(), "%llu %llu", &d, &sr) != 2)
                    return fail(KC_ERR_FORMAT, "%s: unparsable line '%s'", path, tsv_line.c_str());
                if (sr != d) {
                    batch[i++] = {(uint32_t)sr, (uint32_t)d};
                    batch[i++] = {(uint32_t)d, (uint32_t)sr};
                }
            }
        } else {
            for (; i < 100; ++i) {
                std::string l;
                rr.getline(l);
                long long row = 0, col = 0;
                double val;
                int num;
                bool ok = true;
                if (!pattern) {
                    num = std::sscanf(l.c_str(), "%lld %lld %lf", &row, &col, &val);
                    if (array && num == 1) col = row;
                    else if (array || num < 2) ok = false;
                } else {
                    num = std::sscanf(l.c_str(), "%lld %lld", &row, &col);
                    if (array && num == 1) { row = 0; col = 0; }
                    else if (array || num != 2) ok = false;
                }
                if (!ok) break;
                if (col != row) {
                    batch[i] = {(uint32_t)row, (uint32_t)col};
                    if (symmetric) {
                        batch[i] = {(uint32_t)col, (uint32_t)row};
                        ++i;  // leaves batch[i] stale; the for loop increments again
                    }
                }
            }
            // i can end at 101 (symmetric entry on the last slot); resize then
            // appends a value-initialised (0,0) pair, as in the reference.
        }
        batch.resize(i);
        if (i == 0) break;
        for (auto &e : batch) {
            src.push_back(e.first);
            dst.push_back(e.second);
        }
    }
    return KC_OK;
}

}  // namespace kc

extern "C" const char *kc_last_error(void) { return kc::last_error(); }
