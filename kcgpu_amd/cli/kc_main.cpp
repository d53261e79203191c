// kc_main.cpp -- drop-in command line for the reference's `-m kc` path
// (src/main.cu:35-588, flags from include/Config.h:317-363).
//
//   kcgpu -g graph.{mtx,tsv,txt,bel,kco} [-m kc|kcore] [-o degree|degen] [-k K]
//         [-p node|edge] [-q {o|p}<P>[b|n]] [-d dev] [-w x] [-s x] [-x x] [-a ...]
//   kcgpu -m mtx-bel|tsv-bel|tsv-mtx|bel-mtx|txt-bel -g in -r out   (src/main.cu:44-71)
//   kcgpu -m cache -g graph -o degree|degen -r out.kco              (oriented-CSR cache)
//
// Every reference flag takes an argument (getopt "g:r:d:m:x:o:a:k:h:v:s:p:e:q:w:t:");
// -w, -s, -a, -e, -v are accepted for compatibility and do not change the
// result; -t auto lets the autotuner pick algorithm, granularity and P; -x
// prints the graph statistics after orientation; -r names the output of a
// conversion.  A .kco input is
// an oriented-CSR cache file: it is counted with the orientation it was
// saved with, without reading, cleaning or orienting again.  Output keeps the reference's lines (Read graph time,
// Preprocess time, Max Core, the Level/Nodes/Counter line, count time and the
// MOHA ... teps line) with the true count in place of the reference's level
// value, plus a `kcliques=` line and one JSON metrics line.
#include <getopt.h>

#include <algorithm>
#include <cinttypes>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cmath>
#include <cstring>
#include <ctime>
#include <map>
#include <string>
#include <vector>

#include "../../include/kc.h"

namespace {

void stamp(const char *fmt, ...) __attribute__((format(printf, 1, 2)));
void stamp(const char *fmt, ...)
{
    std::time_t t = std::time(nullptr);
    char ts[16];
    std::strftime(ts, sizeof ts, "%H:%M:%S", std::localtime(&t));
    std::printf("[%s] ", ts);
    va_list ap;
    va_start(ap, fmt);
    std::vprintf(fmt, ap);
    va_end(ap);
    std::printf("\n");
}

double now()
{
    timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return t.tv_sec + 1e-9 * t.tv_nsec;
}

void usage()
{
    std::fprintf(stderr,
                 "\nUsage:  kcgpu [options]\n"
                 "    -g <graph file>   .mtx, .tsv, .txt or .bel\n"
                 "    -m <task>         kc (k-clique count, default), lkc (per-vertex counts) or kcore;\n"
                 "                      mtx-bel, tsv-bel, tsv-mtx, bel-mtx, txt-bel convert -g into -r;\n"
                 "                      cache writes the oriented CSR of -g (-o) into -r (.kco)\n"
                 "    -r <file>         output of a conversion or of -m cache\n"
                 "    -o <orient>       degree (default) or degen\n"
                 "    -k <k>            clique size (default 5)\n"
                 "    -p <unit>         node (default) or edge\n"
                 "    -q <spec>         [o|p]<P>[b|n]: orientation or pivoting, partition size\n"
                 "                      P in 1,2,4,...,64, bitmap (b) or index lists (n); default o8b\n"
                 "    -d <device>       HIP device ordinal (default 0)\n"
                 "    -t auto           let the autotuner pick algorithm, granularity and P\n"
                 "    -x <arg>          print graph statistics after orientation\n"
                 "    -w/-s/-a/-e/-v <arg>  accepted for compatibility\n"
                 "    -h <any>          help\n\n");
}

// -x: statistics of the oriented CSR, printed after orientation as the
// reference does (src/main.cu:239-246 on the oriented row pointers with
// nnz = m/2).  Line 1 is MatrixStats (include/utils.cuh:115-184):
// nnz,rows,cols,min,mean,sd,mode,max,count(min),count(mode),count(max) over
// the row lengths, where min/mode/max skip empty rows (the reference calls
// the most frequent length the median and starts max at 1) and mean/sd run
// over every row.  Then PrintMtarixStruct (:187-222): a 10 x 10 grid of the
// share of entries per block, in thousandths ("+" below one, "-" for none).
// Block size n/10 (integer); the reference divides by zero below 10 rows and
// indexes past the grid when n is not a multiple of 10: here the block size
// is at least 1 and the last row and column of blocks take the remainder.
static void print_stats(const kc_oriented *o)
{
    kc_oriented_info oi;
    if (kc_oriented_get_info(o, &oi)) return;
    const uint32_t n = oi.n;
    std::vector<uint32_t> rp(n + 1), col(oi.m ? oi.m : 1), key(n ? n : 1);
    if (kc_oriented_export(o, rp.data(), col.data(), key.data())) return;
    const uint32_t nnz = (uint32_t)oi.m;
    std::map<uint32_t, uint32_t> hist;
    for (uint32_t i = 0; i < n; ++i) ++hist[rp[i + 1] - rp[i]];
    uint32_t mn = 0, mx = 1, mode = 0, cmin = 0, cmode = 0, cmax = 0;
    uint64_t total = 0;
    for (const auto &kv : hist) {
        if (kv.first == 0) continue;
        if (mn == 0) {
            mn = kv.first;
            cmin = kv.second;
        }
        if (cmode < kv.second) {
            mode = kv.first;
            cmode = kv.second;
        }
        total += (uint64_t)kv.second * kv.first;
        mx = kv.first;
        cmax = kv.second;
    }
    const double mean = n ? (double)total / n : 0.0;
    double ss = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const double d = (double)(rp[i + 1] - rp[i]) - mean;
        ss += d * d;
    }
    const double sd = n ? std::sqrt(ss / n) : 0.0;
    std::printf("%u,%u,%u,%u,%.1f,%.1f,%u,%u,%u,%u,%u\n", nnz, n, n, mn, mean, sd, mode, mx, cmin, cmode, cmax);
    const uint32_t unit = std::max<uint32_t>(1, n / 10);
    double st[10][10] = {};
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t ri = std::min<uint32_t>(9, i / unit);
        for (uint32_t e = rp[i]; e < rp[i + 1]; ++e) st[ri][std::min<uint32_t>(9, col[e] / unit)] += 1;
    }
    for (int i = 0; i < 10; ++i) {
        for (int j = 0; j < 10; ++j) {
            const double perc = nnz ? st[i][j] / nnz : 0.0;
            if (perc >= 0.001) std::printf("%d,", (int)(1000 * st[i][j] / nnz));
            else if (perc > 0) std::printf("+,");
            else std::printf("-,");
        }
        std::printf("\n");
    }
}

int die(int rc, const char *what)
{
    std::fprintf(stderr, "error: %s: %s (status %d)\n", what, kc_last_error(), rc);
    return rc == 0 ? 1 : rc;
}

}  // namespace

int main(int argc, char **argv)
{
    const char *graph = nullptr, *dst = nullptr;
    std::string task = "kc", orient = "degree", by = "node", spec = "o8b";
    int k = 5, dev = 0;
    bool autotune = false, stats = false;
    int opt;
    while ((opt = getopt(argc, argv, "g:r:d:m:x:o:a:k:h:v:s:p:e:q:w:t:")) >= 0) {
        switch (opt) {
        case 'g': graph = optarg; break;
        case 'd': dev = std::atoi(optarg); break;
        case 'm': task = optarg; break;
        case 'o': orient = optarg; break;
        case 'k': k = std::atoi(optarg); break;
        case 'p': by = optarg; break;
        case 'q': spec = optarg; break;
        case 't': autotune = std::strcmp(optarg, "auto") == 0; break;
        case 'r': dst = optarg; break;
        case 'x': stats = true; break;
        case 'a': case 'v': case 's': case 'e': case 'w': break;
        case 'h': usage(); return 0;
        default: usage(); return 1;
        }
    }
    if (!graph) {
        usage();
        return 1;
    }
    static const char *conv[] = {"mtx-bel", "tsv-bel", "tsv-mtx", "bel-mtx", "txt-bel"};
    for (int i = 0; i < 5; ++i)
        if (task == conv[i]) {
            if (!dst) {
                std::fprintf(stderr, "error: -m %s needs -r <output file>\n", task.c_str());
                return 1;
            }
            double tc0 = now();
            int crc = kc_convert(i, graph, dst, 0);
            if (crc) return die(crc, task.c_str());
            stamp("Converted %s -> %s (%s) in %f s", graph, dst, task.c_str(), now() - tc0);
            return 0;
        }
    if (task != "kc" && task != "lkc" && task != "kcore" && task != "cache") {
        std::fprintf(stderr, "error: -m %s is not supported (kc, lkc, kcore, cache, mtx-bel, tsv-bel, "
                             "tsv-mtx, bel-mtx, txt-bel)\n", task.c_str());
        return 1;
    }
    if (task == "cache" && !dst) {
        std::fprintf(stderr, "error: -m cache needs -r <output .kco file>\n");
        return 1;
    }
    int omode;
    if (orient == "degree") omode = KC_ORIENT_DEGREE;
    else if (orient == "degen") omode = KC_ORIENT_DEGEN;
    else {
        std::fprintf(stderr, "error: -o %s: the orientation is degree or degen\n", orient.c_str());
        return 1;
    }
    if (task == "kcore") omode = KC_ORIENT_DEGEN;
    kc_query q{};
    q.k = (uint32_t)k;
    q.gran = by == "edge" ? KC_GRAN_EDGE : KC_GRAN_NODE;
    if (by != "edge" && by != "node") {
        std::fprintf(stderr, "error: -p %s: node or edge\n", by.c_str());
        return 1;
    }
    int rc = kc_parse_spec(spec.c_str(), &q);
    if (rc) return die(rc, "-q");
    q.auto_tune = autotune ? 1 : 0;

    std::printf("Welcome ---------------------\n");
    std::printf("    Graph: %s\n    Device Id = %d\n    Main Task = %s\n    Graph Orientation = %s\n"
                "    Process By = %s\n    k: %d\n    KC Config = %s-%u-%s\n",
                graph, dev, task.c_str(), orient.c_str(), by.c_str(), k, q.algo ? "PIV" : "GO", q.part_size,
                q.binary ? "BE" : "NBE");
    if ((rc = kc_select_device(dev))) return die(rc, "device");

    double t0 = now();
    kc_graph *g = nullptr;
    kc_oriented *o = nullptr;
    const size_t glen = std::strlen(graph);
    const bool cached = glen > 4 && std::strcmp(graph + glen - 4, ".kco") == 0;
    if (cached) {
        if ((rc = kc_oriented_load(graph, &o))) return die(rc, "load");
        stamp("Read graph time: %f s", now() - t0);
    } else {
        if ((rc = kc_load(graph, 0, &g))) return die(rc, "load");
        kc_graph_info gi;
        kc_graph_get_info(g, &gi);
        stamp("Read graph time: %f s", now() - t0);
        std::printf("    n = %u, undirected edges = %" PRIu64 ", max degree = %u\n", gi.n, gi.nnz / 2, gi.max_degree);
    }

    double t1 = now();
    if (!cached && (rc = kc_orient(g, omode, &o))) return die(rc, "orient");
    kc_oriented_info oi;
    kc_oriented_get_info(o, &oi);
    if (omode == KC_ORIENT_DEGEN) std::printf("Max Core = %u\n", oi.max_core);
    stamp("Preprocess time: %f s", now() - t1);
    if (stats) print_stats(o);
    if (task == "cache") {
        if ((rc = kc_oriented_save(o, dst))) return die(rc, "cache");
        stamp("Wrote oriented CSR to %s", dst);
    }
    if (task == "kcore" || task == "cache") {
        kc_oriented_free(o);
        kc_graph_free(g);
        return 0;
    }

    kc_result r;
    if (task == "lkc") {
        // src/main.cu:589-748: count, then "count time" and SingleGPU_Kclique_Local::show
        // (kclique/kclique_local.cuh:621-629), which prints the first 50 vertices
        std::vector<uint64_t> local(oi.n ? oi.n : 1);
        double tl = now();
        rc = kc_count_local(o, &q, local.data(), &r);
        double tc = now() - tl;
        if (rc && rc != KC_ERR_OVERFLOW) return die(rc, "count");
        stamp("count time %f s (%f teps)", tc, tc > 0 ? (double)oi.m / tc : 0.0);
        std::printf("Local clique counter for the first %u Nodes:\n", oi.n);
        for (uint32_t i = 0; i < oi.n && i < 50; ++i) std::printf("%u\t%" PRIu64 "\n", i, local[i]);
        std::printf("kcliques=%" PRIu64 " k=%d overflow=%u\n", r.count, k, r.overflow);
        kc_oriented_free(o);
        if (g) kc_graph_free(g);
        return r.overflow ? KC_ERR_OVERFLOW : 0;
    }
    double t2 = now();
    rc = kc_count(o, &q, &r);
    double tc = now() - t2;
    if (rc && rc != KC_ERR_OVERFLOW) return die(rc, "count");
    const unsigned level = (r.gran == KC_GRAN_EDGE) ? (k >= 2 ? k - 2 : 0) : (k >= 1 ? k - 1 : 0);
    if (r.gran == KC_GRAN_NODE)
        std::printf("------------- Level = %u Nodes = %" PRIu64 " Counter = %" PRIu64 "\n", level, r.tasks, r.count);
    else
        std::printf("Level = %u Nodes = %" PRIu64 " Counter = %" PRIu64 "\n", level, r.tasks, r.count);
    stamp("count time %f s", tc);
    stamp("MOHA %" PRIu64 " k-clique (%f teps)", r.count, tc > 0 ? (double)oi.m / tc : 0.0);
    std::printf("kcliques=%" PRIu64 " k=%d overflow=%u\n", r.count, k, r.overflow);
    std::printf("{\"k\": %d, \"count\": %" PRIu64 ", \"overflow\": %u, \"algo\": \"%s\", \"gran\": \"%s\", \"part_size\": %u, "
                "\"n\": %u, \"m_oriented\": %" PRIu64 ", \"preprocess_ms\": %.3f, \"count_ms\": %.3f, "
                "\"kcliques_per_s\": %.6g, \"edges_per_s\": %.6g, \"tasks\": %" PRIu64 "}\n",
                k, r.count, r.overflow, r.algo ? "pivot" : "orient", r.gran ? "edge" : "node", r.part_size, oi.n,
                oi.m, oi.preprocess_ms, r.count_ms, r.count_ms > 0 ? r.count / (r.count_ms / 1e3) : 0.0,
                r.count_ms > 0 ? oi.m / (r.count_ms / 1e3) : 0.0, r.tasks);
    kc_oriented_free(o);
    kc_graph_free(g);
    return r.overflow ? KC_ERR_OVERFLOW : 0;
}
