// Host-only fuzz of the graph loaders and format converters, built with
// AddressSanitizer and UBSan by tests/test_sanitize.py.  Random text, mostly
// valid MTX/TSV files with a few mutated bytes, and random or truncated BEL
// records go through load_file, the reference-order edge reader and every
// kc_convert kind; the sanitizers report any out-of-bounds access, leak or
// undefined arithmetic.  Exit status 0 means no report.
#include <cstdio>
#include <cstdlib>
#include <random>
#include <string>
#include <vector>

#include "../../kcgpu_amd/csrc/kc_internal.h"

static std::mt19937_64 rng(12345);

static std::string random_text(bool mtx)
{
    static const char *toks[] = {"%%MatrixMarket matrix coordinate pattern symmetric\n",
                                 "%%MatrixMarket matrix coordinate real general\n", "% comment\n", "%\n",
                                 "\n", " ", "\t", "0", "1", "2", "3", "17", "4294967295", "4294967296",
                                 "18446744073709551615", "99999999999999999999999", "-1", "-", "+", "1.5",
                                 "e", "x", "\r\n", "  \n"};
    const size_t ntok = sizeof(toks) / sizeof(toks[0]);
    std::string s;
    if (mtx && rng() % 2) s += toks[rng() % 2];
    if (mtx && rng() % 2)
        s += std::to_string(rng() % 20) + " " + std::to_string(rng() % 20) + " " + std::to_string(rng() % 40) + "\n";
    for (int i = 0, n = rng() % 60; i < n; i++) {
        if (rng() % 3 == 0)
            s += std::to_string(rng() % 12) + " " + std::to_string(rng() % 12) + (rng() % 2 ? " 1\n" : "\n");
        else
            s += toks[rng() % ntok];
    }
    return s;
}

static std::string mutated_text(bool mtx)
{
    int n = 1 + rng() % 15, m = rng() % 40;
    std::string s;
    if (mtx)
        s = std::string(rng() % 2 ? "%%MatrixMarket matrix coordinate pattern symmetric\n"
                                  : "%%MatrixMarket matrix coordinate integer general\n") +
            std::to_string(n) + " " + std::to_string(n) + " " + std::to_string(m) + "\n";
    for (int i = 0; i < m; i++)
        s += std::to_string(rng() % (n + 1)) + " " + std::to_string(rng() % (n + 1)) +
             (mtx && rng() % 2 ? "\n" : " 1\n");
    for (int t = rng() % 3; t > 0 && !s.empty(); t--) s[rng() % s.size()] = "0 9\n%-x"[rng() % 7];
    return s;
}

static std::string random_bel(bool records)
{
    std::string s;
    if (records) {
        for (int i = 0, m = rng() % 30; i < 3 * m; i++) {
            uint64_t v = rng() % 20;
            s.append(reinterpret_cast<const char *>(&v), 8);
        }
        if (rng() % 3 == 0 && !s.empty()) s.resize(rng() % s.size());
    } else {
        for (int i = 0, n = rng() % 100; i < n; i++) s.push_back(static_cast<char>(rng() % 4 == 0 ? rng() : rng() % 8));
    }
    return s;
}

int main(int argc, char **argv)
{
    if (argc < 3) {
        fprintf(stderr, "usage: %s <scratch dir> <iterations>\n", argv[0]);
        return 2;
    }
    const std::string dir = argv[1];
    const int iters = atoi(argv[2]);
    const char *exts[] = {".mtx", ".tsv", ".txt", ".bel", ".el"};
    long loaded = 0, rejected = 0;
    for (int it = 0; it < iters; it++) {
        int e = rng() % 5;
        std::string in = dir + "/in" + exts[e];
        std::string s = e == 3 ? random_bel(it % 2) : (it % 2 ? mutated_text(e == 0) : random_text(e == 0));
        FILE *f = fopen(in.c_str(), "wb");
        if (!f) return 2;
        fwrite(s.data(), 1, s.size(), f);
        fclose(f);
        kc::HostCSR g;
        if (kc::load_file(in.c_str(), g) == KC_OK) loaded++; else rejected++;
        std::vector<uint32_t> a, b;
        kc::read_edges_reference(in.c_str(), a, b);
        for (int kind = KC_CONV_MTX_BEL; kind <= KC_CONV_TXT_BEL; kind++)
            kc_convert(kind, in.c_str(), (dir + "/out" + std::to_string(kind)).c_str(), rng() % 2);
    }
    printf("iterations=%d loaded=%ld rejected=%ld\n", iters, loaded, rejected);
    return loaded > 0 && rejected > 0 ? 0 : 1;
}
