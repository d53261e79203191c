// kc_gen.cpp -- seeded synthetic graphs (no reference counterpart: the
// reference reads SNAP files only).  Every generator is deterministic in
// (params, seed): parallel parts use fixed chunks with per-chunk seeds and
// sequential parts run in generation order, so the output does not depend on
// the thread count.
//
// PRNG: xoshiro256** seeded through splitmix64 (Blackman & Vigna).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include <omp.h>

#include "kc_internal.h"

namespace kc {
namespace {

inline uint64_t splitmix64(uint64_t &x)
{
    uint64_t z = (x += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

struct Rng {
    uint64_t s[4];
    explicit Rng(uint64_t seed)
    {
        uint64_t x = seed;
        for (auto &w : s) w = splitmix64(x);
    }
    static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    inline uint64_t next()
    {
        uint64_t r = rotl(s[1] * 5, 7) * 9;
        uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    // uniform in [0, n) without modulo bias (Lemire)
    inline uint64_t below(uint64_t n)
    {
        for (;;) {
            unsigned __int128 m = (unsigned __int128)next() * n;
            uint64_t lo = (uint64_t)m;
            if (lo >= n || lo >= (uint64_t)(-n) % n) return (uint64_t)(m >> 64);
        }
    }
    inline double unit() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

// open-addressing set of undirected edges (key = lo << 32 | hi, never 0 since lo < hi)
struct EdgeSet {
    std::vector<uint64_t> t;
    uint64_t mask;
    explicit EdgeSet(uint64_t want)
    {
        uint64_t cap = 16;
        while (cap < want * 2) cap <<= 1;
        t.assign(cap, 0);
        mask = cap - 1;
    }
    bool insert(uint32_t a, uint32_t b)
    {
        if (a > b) std::swap(a, b);
        uint64_t key = ((uint64_t)a << 32) | b;
        uint64_t h = key * 0x9e3779b97f4a7c15ull;
        uint64_t i = (h ^ (h >> 29)) & mask;
        for (;;) {
            if (t[i] == key) return false;
            if (t[i] == 0) { t[i] = key; return true; }
            i = (i + 1) & mask;
        }
    }
};

int gen_er(uint32_t n, uint64_t m, uint64_t seed, std::vector<uint32_t> &s, std::vector<uint32_t> &d)
{
    if (n < 2) return m ? fail(KC_ERR_ARG, "ER: n=%u too small", n) : KC_OK;
    uint64_t maxm = (uint64_t)n * (n - 1) / 2;
    if (m > maxm) return fail(KC_ERR_ARG, "ER: m=%llu > n(n-1)/2", (unsigned long long)m);
    Rng r(seed);
    EdgeSet es(m);
    s.reserve(m);
    d.reserve(m);
    while (s.size() < m) {
        uint32_t a = (uint32_t)r.below(n), b = (uint32_t)r.below(n);
        if (a == b || !es.insert(a, b)) continue;
        s.push_back(a);
        d.push_back(b);
    }
    return KC_OK;
}

// Chung-Lu: w_i = (i + i0)^(-1/(gamma-1)); i0 solves max expected degree
// c * i0^(-alpha) = max_degree with c = 2m / sum_i w_i.  Endpoints are drawn
// proportional to w through an alias table; loops and repeated pairs are
// rejected until m distinct edges exist.
double weight_sum(uint32_t n, double i0, double alpha)
{
    const uint32_t CH = 1 << 16;
    uint32_t nch = (n + CH - 1) / CH;
    std::vector<double> part(nch, 0.0);
#pragma omp parallel for schedule(static)
    for (uint32_t c = 0; c < nch; ++c) {
        double acc = 0;
        uint32_t e = std::min<uint64_t>((uint64_t)(c + 1) * CH, n);
        for (uint32_t i = c * CH; i < e; ++i) acc += std::pow((double)i + i0, -alpha);
        part[c] = acc;
    }
    double t = 0;
    for (double x : part) t += x;  // fixed order
    return t;
}

int gen_chung_lu(uint32_t n, uint64_t m, double gamma, double maxdeg, uint64_t seed,
                 std::vector<uint32_t> &s, std::vector<uint32_t> &d)
{
    if (n < 2 || gamma <= 1.0 || maxdeg <= 0) return fail(KC_ERR_ARG, "Chung-Lu: bad parameters");
    uint64_t maxm = (uint64_t)n * (n - 1) / 2;
    if (m > maxm / 2) return fail(KC_ERR_ARG, "Chung-Lu: m too large for n");
    double alpha = 1.0 / (gamma - 1.0);
    // f(i0) = 2m i0^-alpha / S(i0) decreases in i0
    double lo = 1e-3, hi = (double)n;
    for (int it = 0; it < 60; ++it) {
        double mid = std::sqrt(lo * hi);
        double f = 2.0 * (double)m * std::pow(mid, -alpha) / weight_sum(n, mid, alpha);
        if (f > maxdeg) lo = mid; else hi = mid;
        if (hi / lo < 1.0 + 1e-9) break;
    }
    double i0 = std::sqrt(lo * hi);
    // alias table (Vose), built sequentially
    std::vector<double> w(n);
    double tot = 0;
    for (uint32_t i = 0; i < n; ++i) { w[i] = std::pow((double)i + i0, -alpha); tot += w[i]; }
    std::vector<double> prob(n);
    std::vector<uint32_t> alias(n);
    std::vector<uint32_t> small, large;
    small.reserve(n);
    large.reserve(n);
    for (uint32_t i = 0; i < n; ++i) {
        prob[i] = w[i] * n / tot;
        (prob[i] < 1.0 ? small : large).push_back(i);
    }
    while (!small.empty() && !large.empty()) {
        uint32_t a = small.back(); small.pop_back();
        uint32_t b = large.back(); large.pop_back();
        alias[a] = b;
        prob[b] = prob[b] + prob[a] - 1.0;
        (prob[b] < 1.0 ? small : large).push_back(b);
    }
    for (uint32_t i : large) { prob[i] = 1.0; alias[i] = i; }
    for (uint32_t i : small) { prob[i] = 1.0; alias[i] = i; }
    Rng r(seed);
    auto draw = [&]() -> uint32_t {
        uint32_t i = (uint32_t)r.below(n);
        return r.unit() < prob[i] ? i : alias[i];
    };
    EdgeSet es(m);
    s.reserve(m);
    d.reserve(m);
    uint64_t tries = 0, limit = 64 * m + 1000000;
    while (s.size() < m) {
        if (++tries > limit) return fail(KC_ERR_ARG, "Chung-Lu: cannot reach m distinct edges");
        uint32_t a = draw(), b = draw();
        if (a == b || !es.insert(a, b)) continue;
        s.push_back(a);
        d.push_back(b);
    }
    return KC_OK;
}

// R-MAT (Chakrabarti et al.): each of 2^scale * edge_factor samples descends
// `scale` levels choosing quadrant a/b/c/d.  Chunks of 2^16 samples use their
// own stream seeded from (seed, chunk), so the result is thread-count free.
int gen_rmat(uint32_t scale, uint32_t ef, double a, double b, double c, uint64_t seed,
             std::vector<uint32_t> &s, std::vector<uint32_t> &d)
{
    if (scale == 0 || scale > 31 || ef == 0) return fail(KC_ERR_ARG, "R-MAT: bad scale/edge factor");
    if (a < 0 || b < 0 || c < 0 || a + b + c > 1.0) return fail(KC_ERR_ARG, "R-MAT: bad probabilities");
    uint64_t ns = (uint64_t)ef << scale;
    s.resize(ns);
    d.resize(ns);
    const double t1 = a, t2 = a + b, t3 = a + b + c;
    const uint64_t T1 = (uint64_t)(t1 * 4294967296.0), T2 = (uint64_t)(t2 * 4294967296.0),
                   T3 = (uint64_t)(t3 * 4294967296.0);
    const uint64_t CH = 1 << 16;
    uint64_t nch = (ns + CH - 1) / CH;
#pragma omp parallel for schedule(dynamic, 4)
    for (uint64_t ch = 0; ch < nch; ++ch) {
        uint64_t x = seed ^ (ch * 0xd1b54a32d192ed03ull);
        Rng r(splitmix64(x));
        uint64_t e = std::min(ns, (ch + 1) * CH);
        for (uint64_t i = ch * CH; i < e; ++i) {
            uint32_t u = 0, v = 0;
            uint64_t bits = 0;
            int avail = 0;
            for (uint32_t l = 0; l < scale; ++l) {
                if (avail == 0) { bits = r.next(); avail = 2; }
                uint64_t q = bits & 0xFFFFFFFFull;
                bits >>= 32;
                --avail;
                uint32_t bu = q >= T2, bv = (q >= T1 && q < T2) || q >= T3;
                u = (u << 1) | bu;
                v = (v << 1) | bv;
            }
            s[i] = u;
            d[i] = v;
        }
    }
    return KC_OK;
}

// Barabasi-Albert: a star on attach+1 vertices, then every new vertex links to
// `attach` distinct vertices drawn uniformly from the list in which each
// vertex appears once per incident edge.
int gen_ba(uint32_t n, uint32_t attach, uint64_t seed, std::vector<uint32_t> &s, std::vector<uint32_t> &d)
{
    if (attach == 0 || n <= attach) return fail(KC_ERR_ARG, "BA: need n > attach >= 1");
    Rng r(seed);
    std::vector<uint32_t> rep;
    rep.reserve(2ull * n * attach);
    for (uint32_t i = 1; i <= attach; ++i) {
        s.push_back(0);
        d.push_back(i);
        rep.push_back(0);
        rep.push_back(i);
    }
    std::vector<uint32_t> tg(attach);
    for (uint32_t v = attach + 1; v < n; ++v) {
        uint32_t got = 0;
        while (got < attach) {
            uint32_t t = rep[r.below(rep.size())];
            bool dup = false;
            for (uint32_t i = 0; i < got; ++i) dup |= (tg[i] == t);
            if (!dup) tg[got++] = t;
        }
        for (uint32_t i = 0; i < attach; ++i) {
            s.push_back(v);
            d.push_back(tg[i]);
            rep.push_back(tg[i]);
            rep.push_back(v);
        }
    }
    return KC_OK;
}

}  // namespace

int generate(const kc_gen_params &p, std::vector<uint32_t> &src, std::vector<uint32_t> &dst, uint32_t &n)
{
    src.clear();
    dst.clear();
    switch (p.kind) {
    case KC_GEN_ER:
        n = p.n;
        return gen_er(p.n, p.m, p.seed, src, dst);
    case KC_GEN_CHUNG_LU:
        n = p.n;
        return gen_chung_lu(p.n, p.m, p.gamma, p.max_degree, p.seed, src, dst);
    case KC_GEN_RMAT:
        n = p.scale && p.scale <= 31 ? (1u << p.scale) : 0;
        return gen_rmat(p.scale, p.edge_factor, p.a, p.b, p.c, p.seed, src, dst);
    case KC_GEN_BA:
        n = p.n;
        return gen_ba(p.n, p.attach, p.seed, src, dst);
    default:
        return fail(KC_ERR_ARG, "unknown generator kind %d", p.kind);
    }
}

}  // namespace kc
