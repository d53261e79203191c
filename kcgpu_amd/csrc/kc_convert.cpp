// kc_convert.cpp -- on-disk format converters (-m mtx-bel, tsv-bel, tsv-mtx,
// bel-mtx, txt-bel) and the binary oriented-CSR cache format.
//
// The converters follow graph::MtB_Writer (include/FIleReader.cuh:480-933)
// record for record:
//   BEL output   24-byte records, u64 dst @0, u64 src @8, u64 0 @16
//   MTX output   "%%MatrixMarket matrix coordinate general", "n n m", then
//                "src dst weight" per edge, n = src of the last edge
// Where the reference has undefined behaviour (reads past a missing size line,
// fileEdges[m-1] on an empty input, an 8-byte fscanf into a 4-byte int, a
// 100-byte line buffer, an endless loop on an empty line) this code returns an
// error or takes the defined reading stated at that spot.
#include <algorithm>
#include <cerrno>
#include <cinttypes>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "kc_internal.h"

namespace kc {
namespace {

using Pair = std::pair<uint64_t, uint64_t>;

int slurp(const char *path, std::string &buf)
{
    FILE *f = std::fopen(path, "rb");
    if (!f) return fail(KC_ERR_IO, "cannot open '%s': %s", path, std::strerror(errno));
    std::fseek(f, 0, SEEK_END);
    long len = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    if (len < 0) {
        std::fclose(f);
        return fail(KC_ERR_IO, "cannot size '%s'", path);
    }
    buf.resize((size_t)len);
    size_t got = len ? std::fread(&buf[0], 1, (size_t)len, f) : 0;
    std::fclose(f);
    if (got != (size_t)len) return fail(KC_ERR_IO, "short read on '%s'", path);
    return KC_OK;
}

// Whitespace-separated token reader with fscanf("%llu") semantics: leading
// white space (newlines included) is skipped, an optional sign and decimal
// digits form the number; false when no digits follow (or at end of input).
// A '-' sign is refused as a format error through *neg (fscanf would wrap it).
bool scan_u64(const char *&p, const char *end, uint64_t &v, bool &neg)
{
    while (p < end && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r' || *p == '\v' || *p == '\f')) ++p;
    neg = false;
    const char *q = p;
    if (q < end && (*q == '+' || *q == '-')) {
        neg = *q == '-';
        ++q;
    }
    if (q >= end || *q < '0' || *q > '9') return false;
    uint64_t x = 0;
    bool sat = false;
    while (q < end && *q >= '0' && *q <= '9') {
        uint64_t d = (uint64_t)(*q - '0');
        if (x > (UINT64_MAX - d) / 10) sat = true;
        else x = x * 10 + d;
        ++q;
    }
    v = sat ? UINT64_MAX : x;  // strtoull saturates
    p = q;
    return true;
}

// Triples "a b c" read as fscanf("%llu %llu %llu") does in a do/while loop
// (include/FIleReader.cuh:669-676, :721-730): stop at the first triple that
// does not scan completely.
int scan_triples(const char *path, const std::string &buf, std::vector<uint64_t> &out)
{
    const char *p = buf.data(), *end = p + buf.size();
    for (;;) {
        uint64_t t[3];
        bool neg;
        int got = 0;
        while (got < 3 && scan_u64(p, end, t[got], neg)) {
            if (neg) return fail(KC_ERR_FORMAT, "%s: negative vertex id or weight", path);
            ++got;
        }
        if (got < 3) break;
        out.insert(out.end(), t, t + 3);
    }
    return KC_OK;
}

int write_bel(const char *path, const std::vector<Pair> &edges)
{
    FILE *f = std::fopen(path, "wb");
    if (!f) return fail(KC_ERR_IO, "cannot create '%s': %s", path, std::strerror(errno));
    const size_t shard = 100000;
    std::vector<uint64_t> rec(3 * shard);
    for (size_t i = 0; i < edges.size(); i += shard) {
        size_t e = std::min(edges.size(), i + shard), c = 0;
        for (size_t j = i; j < e; ++j) {
            rec[c++] = edges[j].second;  // dst
            rec[c++] = edges[j].first;   // src
            rec[c++] = 0;                // weight
        }
        if (std::fwrite(rec.data(), 8, c, f) != c) {
            std::fclose(f);
            return fail(KC_ERR_IO, "write to '%s' failed", path);
        }
    }
    if (std::fclose(f) != 0) return fail(KC_ERR_IO, "closing '%s' failed", path);
    return KC_OK;
}

// MtB_Writer's Matrix Market writer (include/FIleReader.cuh:738-750, :795-807)
// with T = uint (ids kept to 32 bits) and ValueT = int.
int write_mtx(const char *path, const std::vector<Pair> &edges, const std::vector<int32_t> &w)
{
    if (edges.empty()) return fail(KC_ERR_FORMAT, "no edges to write to '%s'", path);
    FILE *f = std::fopen(path, "w");
    if (!f) return fail(KC_ERR_IO, "cannot create '%s': %s", path, std::strerror(errno));
    const uint64_t n = edges.back().first;
    std::fprintf(f, "%%%%MatrixMarket matrix coordinate general\n%" PRIu64 " %" PRIu64 " %zu\n", n, n, edges.size());
    for (size_t i = 0; i < edges.size(); ++i)
        std::fprintf(f, "%" PRIu64 " %" PRIu64 " %d\n", edges[i].first, edges[i].second, w[i]);
    if (std::fclose(f) != 0) return fail(KC_ERR_IO, "closing '%s' failed", path);
    return KC_OK;
}

int check_u32(const char *path, uint64_t a, uint64_t b)
{
    if (a > 0xFFFFFFFFull || b > 0xFFFFFFFFull)
        return fail(KC_ERR_FORMAT, "%s: vertex id does not fit 32 bits", path);
    return KC_OK;
}

// mtx-bel: write_market_bel<uint,int>(src, dst, makeFull=false)
// (include/FIleReader.cuh:484-656, called from src/main.cu:48-50).
int mtx_to_bel(const char *in, const char *out, unsigned flags)
{
    std::string buf;
    int rc = slurp(in, buf);
    if (rc) return rc;
    const char *p = buf.data(), *end = p + buf.size();
    bool symmetric = false, pattern = false, array = false, banner = false;
    size_t lineno = 0;
    auto next_line = [&](const char *&b, const char *&e) {
        if (p >= end) return false;
        b = p;
        e = (const char *)std::memchr(p, '\n', (size_t)(end - p));
        if (!e) e = end;
        p = e < end ? e + 1 : end;
        ++lineno;
        return true;
    };
    const char *b = nullptr, *e = nullptr;
    for (;;) {
        if (!next_line(b, e)) return fail(KC_ERR_FORMAT, "%s: no Matrix Market size line", in);
        if (b < e && *b == '%') {
            // The reference re-reads the flags from every '%' line of two or
            // more characters, so a later comment clears them; the default
            // here reads them from the first "%%" banner only.
            bool take = (flags & KC_CONV_REFERENCE_FLAGS) ? (e - b >= 2) : (!banner && e - b >= 2 && b[1] == '%');
            if (take) {
                std::string l(b, e);
                symmetric = l.find("symmetric") != std::string::npos;
                pattern = l.find("pattern") != std::string::npos;
                array = l.find("array") != std::string::npos;
                banner = true;
            }
            continue;
        }
        break;
    }
    uint64_t r, c, nz;
    bool neg;
    const char *q = b;
    if (!scan_u64(q, e, r, neg) || neg || !scan_u64(q, e, c, neg) || neg || !scan_u64(q, e, nz, neg) || neg)
        return fail(KC_ERR_FORMAT, "%s:%zu: size line needs rows cols entries", in, lineno);
    if (array) return fail(KC_ERR_FORMAT, "%s: dense 'array' Matrix Market is not a graph", in);
    if (r != c)
        return fail(KC_ERR_FORMAT, "%s: matrix is not square (%" PRIu64 " x %" PRIu64 ")", in, r, c);
    std::vector<Pair> edges;
    edges.reserve(symmetric ? 2 * nz : nz);
    for (uint64_t i = 0; i < nz; ++i) {
        if (!next_line(b, e)) return fail(KC_ERR_FORMAT, "%s: %" PRIu64 " entries announced, %" PRIu64 " found", in, nz, i);
        q = b;
        uint64_t row, col;
        bool n1, n2;
        // a value column, if any, is parsed by the reference and then dropped
        // (BEL weights are written as 0)
        if (!scan_u64(q, e, row, n1) || !scan_u64(q, e, col, n2) || n1 || n2)
            return fail(KC_ERR_FORMAT, "%s:%zu: badly formed edge", in, lineno);
        (void)pattern;
        if ((rc = check_u32(in, row, col))) return rc;
        edges.emplace_back(row, col);
        if (symmetric && row != col) edges.emplace_back(col, row);
    }
    std::sort(edges.begin(), edges.end());
    return write_bel(out, edges);
}

// tsv-bel: write_tsv_bel<uint64,uint64> (include/FIleReader.cuh:659-706):
// "dst src weight" triples in file order, no sort.
int tsv_to_bel(const char *in, const char *out)
{
    std::string buf;
    std::vector<uint64_t> t;
    int rc = slurp(in, buf);
    if (rc || (rc = scan_triples(in, buf, t))) return rc;
    std::vector<Pair> edges(t.size() / 3);
    for (size_t i = 0; i < edges.size(); ++i) edges[i] = Pair(t[3 * i + 1], t[3 * i]);
    return write_bel(out, edges);
}

// tsv-mtx: write_tsv_market<uint,int> (include/FIleReader.cuh:710-752).  The
// reference scans the weight with %llu into an int; its low 32 bits are kept.
int tsv_to_mtx(const char *in, const char *out)
{
    std::string buf;
    std::vector<uint64_t> t;
    int rc = slurp(in, buf);
    if (rc || (rc = scan_triples(in, buf, t))) return rc;
    size_t m = t.size() / 3;
    std::vector<Pair> edges(m);
    std::vector<int32_t> w(m);
    for (size_t i = 0; i < m; ++i) {
        if ((rc = check_u32(in, t[3 * i], t[3 * i + 1]))) return rc;
        edges[i] = Pair(t[3 * i + 1], t[3 * i]);
        w[i] = (int32_t)(uint32_t)t[3 * i + 2];
    }
    return write_mtx(out, edges, w);
}

// bel-mtx: write_bel_market<uint,int> (include/FIleReader.cuh:756-810): whole
// 24-byte records only; the weight is the low 32 bits of the u64 at byte 16.
int bel_to_mtx(const char *in, const char *out)
{
    std::string buf;
    int rc = slurp(in, buf);
    if (rc) return rc;
    size_t m = buf.size() / 24;
    std::vector<Pair> edges(m);
    std::vector<int32_t> w(m);
    for (size_t i = 0; i < m; ++i) {
        uint64_t dst, src, wt;
        std::memcpy(&dst, buf.data() + 24 * i, 8);
        std::memcpy(&src, buf.data() + 24 * i + 8, 8);
        std::memcpy(&wt, buf.data() + 24 * i + 16, 8);
        if ((rc = check_u32(in, src, dst))) return rc;
        edges[i] = Pair(src, dst);
        w[i] = (int32_t)(uint32_t)wt;
    }
    return write_mtx(out, edges, w);
}

// txt-bel: write_txt_bel<uint,uint>(src, dst, makeFull=true, numCol=2,
// srcIndex=0) (include/FIleReader.cuh:814-932, src/main.cu:68-70): lines
// starting with '%' are skipped, "a b" with a != b gives (a,b) and (b,a);
// sorted and deduplicated.  Empty lines (an endless loop in the reference)
// are skipped; lines of any length are read.
int txt_to_bel(const char *in, const char *out)
{
    std::string buf;
    int rc = slurp(in, buf);
    if (rc) return rc;
    std::vector<Pair> edges;
    const char *p = buf.data(), *end = p + buf.size();
    while (p < end) {
        const char *e = (const char *)std::memchr(p, '\n', (size_t)(end - p));
        if (!e) e = end;
        const char *q = p;
        p = e < end ? e + 1 : end;
        if (q == e || *q == '%') continue;
        uint64_t a, b;
        bool n1, n2;
        // sscanf("%llu %llu") on the line: the second number may sit on the
        // same line only
        if (!scan_u64(q, e, a, n1) || !scan_u64(q, e, b, n2)) continue;
        if (n1 || n2) return fail(KC_ERR_FORMAT, "%s: negative vertex id", in);
        if (a == b) continue;
        if ((rc = check_u32(in, a, b))) return rc;
        edges.emplace_back(a, b);
        edges.emplace_back(b, a);
    }
    std::sort(edges.begin(), edges.end());
    edges.erase(std::unique(edges.begin(), edges.end()), edges.end());
    return write_bel(out, edges);
}

}  // namespace
}  // namespace kc

extern "C" int kc_convert(int kind, const char *src_path, const char *dst_path, unsigned flags)
{
    using namespace kc;
    if (!src_path || !dst_path) return fail(KC_ERR_ARG, "kc_convert: null path");
    switch (kind) {
    case KC_CONV_MTX_BEL: return mtx_to_bel(src_path, dst_path, flags);
    case KC_CONV_TSV_BEL: return tsv_to_bel(src_path, dst_path);
    case KC_CONV_TSV_MTX: return tsv_to_mtx(src_path, dst_path);
    case KC_CONV_BEL_MTX: return bel_to_mtx(src_path, dst_path);
    case KC_CONV_TXT_BEL: return txt_to_bel(src_path, dst_path);
    default: return fail(KC_ERR_ARG, "kc_convert: unknown conversion %d", kind);
    }
}
