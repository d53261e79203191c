/* Built with AddressSanitizer and UBSan by tests/test_sanitize.py: random
 * graphs of up to 40 vertices (edge density 0..1) go through the oracle's peel,
 * both orientations, the four traversals and the per-vertex counts at k = 1..9;
 * every count must equal the brute-force count (per-vertex sums k times it).
 * Exit status 0 means no mismatch and no sanitizer report. */
#include <stdio.h>
#include <stdlib.h>
#include <stdint.h>
typedef uint64_t u64; typedef uint32_t u32;
u64 orc_binom(u32 n, u32 r, int *ovf);
int64_t orc_peel_priority(u32 n, const u64 *rowptr, const u32 *col, u32 *prio);
u64 orc_orient(u32 n, const u64 *rowptr, const u32 *col, const u32 *key, u64 *out_rowptr, u32 *out_col);
int orc_count(u32 n, const u64 *rowptr, const u32 *col, u32 k, int algo, int gran, u64 stride, u64 offset,
              int nthreads, u64 *count, int *overflow, u64 *tasks_done, u64 *alg_bytes);
int orc_count_local(u32 n, const u64 *rowptr, const u32 *col, u32 k, int algo, int gran, int nthreads,
                    u64 *local, int *overflow);
u64 orc_brute_force(u32 n, const u64 *adj, u32 k, int *overflow);
static u64 s = 88172645463325252ull;
static u64 rnd(void) { s ^= s << 13; s ^= s >> 7; s ^= s << 17; return s; }
int main(int argc, char **argv) {
    int iters = argc > 1 ? atoi(argv[1]) : 300, bad = 0;
    for (int it = 0; it < iters; it++) {
        u32 n = 1 + rnd() % 40;
        u64 adj[64] = {0};
        double p = (rnd() % 100) / 100.0;
        for (u32 u = 0; u < n; u++) for (u32 v = u + 1; v < n; v++)
            if ((rnd() % 1000) < p * 1000) { adj[u] |= 1ull << v; adj[v] |= 1ull << u; }
        u64 *rp = malloc(8 * (n + 1)); u32 *col = malloc(4 * n * n + 4);
        u64 w = 0;
        for (u32 u = 0; u < n; u++) { rp[u] = w; for (u32 v = 0; v < n; v++) if (adj[u] >> v & 1) col[w++] = v; }
        rp[n] = w;
        u32 *prio = malloc(4 * n);
        orc_peel_priority(n, rp, col, prio);
        for (int key = 0; key < 2; key++) {
            u64 *orp = malloc(8 * (n + 1)); u32 *oc = malloc(4 * (w / 2) + 4);
            orc_orient(n, rp, col, key ? prio : NULL, orp, oc);
            for (u32 k = 1; k <= 9; k++) {
                int o1; u64 bf = orc_brute_force(n, adj, k, &o1);
                for (int algo = 0; algo < 2; algo++) for (int gran = 0; gran < 2; gran++) {
                    u64 c, td, ab; int ov;
                    orc_count(n, orp, oc, k, algo, gran, 1, 0, 2, &c, &ov, &td, &ab);
                    if (c != bf) { bad++; printf("count mismatch n=%u k=%u algo=%d gran=%d %llu vs %llu\n", n, k, algo, gran, (unsigned long long)c, (unsigned long long)bf); }
                    u64 *loc = calloc(n, 8);
                    orc_count_local(n, orp, oc, k, algo, gran, 2, loc, &ov);
                    u64 sum = 0; for (u32 v = 0; v < n; v++) sum += loc[v];
                    if (sum != k * bf) { bad++; printf("local mismatch n=%u k=%u\n", n, k); }
                    free(loc);
                }
            }
            free(orp); free(oc);
        }
        free(rp); free(col); free(prio);
    }
    printf("iters=%d bad=%d\n", iters, bad);
    return bad != 0;
}
