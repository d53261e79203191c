/*
 * kc_oracle.c -- CPU restatement of the KCGPU k-clique path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/,
 * __graft_entry__.smoke() and the cpu_baseline leg of bench.py load it.
 * The product library (kcgpu_amd/csrc) never links or calls it.
 *
 * What is restated here (reference file:line in each function comment):
 *   - orc_peel_priority : level-synchronous k-core peel that produces the
 *                         degeneracy key used by `-o degen`
 *                         (kcore/kcore.cuh:442-584, :160-230, :23-47,
 *                          truss/kernels.cuh:124-174)
 *   - orc_orient        : keep-flag rule + stable compaction + row histogram
 *                         + exclusive scan that turn the symmetric CSR into
 *                         the oriented CSR (include/main_support.cuh:54-128,
 *                         :133-167; src/main.cu:179-209)
 *   - orc_count         : node/edge x orientation/pivoting counters
 *                         (kclique/kckernels_orientation.cuh:3-44,319-650,
 *                          kclique/kckernels_pivoting.cuh:2-346,
 *                          kclique/kckernels.cuh:2263-2630), with the
 *                         task filters of kclique/kclique.cuh:1609-1614,
 *                         :2344-2355
 *   - orc_count_local   : per-vertex counts of `-m lkc`
 *                         (kclique/kclique_local.cuh:24-639 with the kernels
 *                          of kclique/kckernels_local.cuh)
 *   - orc_binom         : exact binomial coefficients that replace the
 *                         double-built nCr table (kclique/kclique_local.cuh:345-361)
 *   - orc_brute_force   : all-k-subsets counter for tiny graphs (n <= 64);
 *                         not a reference function, it is the mathematical
 *                         definition used to pin the counters above.
 *
 * Deliberate differences from the reference (all of them make the count the
 * mathematical k-clique count of the simple graph):
 *   - no degree-window buckets (the 1024-wide windows truncate pivot runs,
 *     kclique/kclique.cuh:1725,1790,2331,2400);
 *   - edge pivoting does not skip tasks whose common-neighbour set has no
 *     internal edge (kclique/kckernels.cuh:2375-2376), so k=3 is exact;
 *   - no MAXDEPTH stack limit (kclique/kclique.cuh:1664);
 *   - the sum is taken in 128-bit integers and an overflow of the uint64
 *     result is reported instead of wrapping (the reference atomicAdds wrap).
 *   - bitmaps use 64-bit words and the SCT children exclude earlier branch
 *     vertices in ascending index order; the reference uses 32-bit words and
 *     a block-wise order (kclique/kckernels_pivoting.cuh:197,237).  Any total
 *     order over the branch set yields the same count.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef unsigned __int128 u128;
typedef uint64_t u64;
typedef uint32_t u32;

/* ------------------------------------------------------------------------ */
/* Exact binomial coefficients.                                              */
/* The reference fills nCr[n*401+r] from a double Pascal triangle            */
/* (kclique/kclique_local.cuh:345-361) and reads it from nCr.txt at run time  */
/* (kclique/kclique.cuh:1694-1717).  Here: Pascal's rule in 128 bits with a   */
/* saturation flag; entries above 2^64-1 are flagged.                        */
/* ------------------------------------------------------------------------ */
#define BIG ((u128)1 << 127)

static u128 binom128(u32 n, u32 r)
{
    /* C(n,r) by the multiplicative formula, saturating at BIG. */
    if (r > n) return 0;
    if (r > n - r) r = n - r;
    u128 c = 1;
    for (u32 i = 1; i <= r; ++i) {
        /* c = c * (n - r + i) / i  is exact at every step */
        u128 num = (u128)(n - r + i);
        if (c > BIG / num) return BIG;
        c = c * num / i;
    }
    return c;
}

/* orc_binom: value of C(n, r) as uint64; *ovf = 1 when it does not fit. */
u64 orc_binom(u32 n, u32 r, int *ovf)
{
    u128 c = binom128(n, r);
    *ovf = (c > (u128)UINT64_MAX);
    return (u64)c;
}

/* ------------------------------------------------------------------------ */
/* Degeneracy peel (kcore/kcore.cuh:442-584).                                */
/*  - deg[v] starts at the full degree (getNodeDegree_kernel, :50-59);       */
/*  - priority starts at numEdges of the full CSR (:471-472);                 */
/*  - level L takes the unpeeled vertices whose degree is exactly L           */
/*    (bucket_scan / filter_with_random_append, :321-362,                     */
/*     truss/kernels.cuh:160-174);                                            */
/*  - every round: if no vertex is left after it the loop breaks before the   */
/*    round is assigned (:495-498); otherwise every frontier vertex gets      */
/*    priority = round (update_priority / kernel_partition_level_next :192)   */
/*    and each neighbour outside the frontier loses one degree; a neighbour   */
/*    whose degree was L+1 joins the next round (process_degree, :23-47).     */
/*  - round++ per round (:564); level++ when a level drains (:569);           */
/*    "Max Core" = level - 1 after the loop (:581-583).                       */
/* Decrements of already-peeled vertices (the reference's stale marks) can    */
/* never re-queue them, so they are skipped here.                             */
/* Returns the max core.                                                      */
/* ------------------------------------------------------------------------ */
int64_t orc_peel_priority(u32 n, const u64 *rowptr, const u32 *col, u32 *prio)
{
    u64 nnz = rowptr[n];
    int64_t *deg = (int64_t *)malloc(sizeof(int64_t) * (n ? n : 1));
    unsigned char *state = (unsigned char *)calloc(n ? n : 1, 1); /* 0 live, 1 frontier, 2 peeled */
    u32 *cur = (u32 *)malloc(sizeof(u32) * (n ? n : 1));
    u32 *nxt = (u32 *)malloc(sizeof(u32) * (n ? n : 1));
    for (u32 v = 0; v < n; ++v) {
        deg[v] = (int64_t)(rowptr[v + 1] - rowptr[v]);
        prio[v] = (u32)nnz;
    }
    int64_t todo = n;
    int64_t level = 0;
    u32 round = 0;
    while (todo > 0) {
        size_t nc = 0;
        for (u32 v = 0; v < n; ++v)
            if (state[v] == 0 && deg[v] == level) cur[nc++] = v;
        int done = 0;
        while (nc > 0) {
            todo -= (int64_t)nc;
            if (todo == 0) { done = 1; break; }
            for (size_t i = 0; i < nc; ++i) { state[cur[i]] = 1; prio[cur[i]] = round; }
            size_t nn = 0;
            for (size_t i = 0; i < nc; ++i) {
                u32 v = cur[i];
                for (u64 e = rowptr[v]; e < rowptr[v + 1]; ++e) {
                    u32 x = col[e];
                    if (state[x] != 0) continue;
                    int64_t old = deg[x]--;
                    if (old == level + 1) nxt[nn++] = x;
                }
            }
            for (size_t i = 0; i < nc; ++i) state[cur[i]] = 2;
            u32 *t = cur; cur = nxt; nxt = t;
            nc = nn;
            round++;
        }
        level++;
        if (done) break;
    }
    free(deg); free(state); free(cur); free(nxt);
    return level - 1;
}

/* ------------------------------------------------------------------------ */
/* Orientation (src/main.cu:179-209).                                        */
/* keep[e] for full-CSR entry e=(src,dst) (include/main_support.cuh):        */
/*   degree  (4-arg init, :54-86): key = full degree; the dstLen<srcLen arm   */
/*           can never fire when keys are degrees, so the rule is             */
/*           (deg src, src) < (deg dst, dst)                                 */
/*   degen   (5-arg init, :89-128): (prio src, src) < (prio dst, dst)        */
/* Stable compaction (CUBSelect, src/main.cu:192-193) keeps the row order;    */
/* the per-row kept count (warp_detect_deleted_edges, :133-167) is scanned    */
/* exclusively into the new row pointer (src/main.cu:200-202).               */
/* key == NULL selects the degree key.                                        */
/* out_rowptr has n+1 entries; out_col has room for nnz/2 entries (symmetric  */
/* simple input).  Returns the number of oriented edges written.             */
/* ------------------------------------------------------------------------ */
u64 orc_orient(u32 n, const u64 *rowptr, const u32 *col, const u32 *key,
               u64 *out_rowptr, u32 *out_col)
{
    u64 w = 0;
    for (u32 u = 0; u < n; ++u) {
        out_rowptr[u] = w;
        u64 ku = key ? key[u] : (rowptr[u + 1] - rowptr[u]);
        for (u64 e = rowptr[u]; e < rowptr[u + 1]; ++e) {
            u32 v = col[e];
            u64 kv = key ? key[v] : (rowptr[v + 1] - rowptr[v]);
            if (ku < kv || (ku == kv && u < v)) out_col[w++] = v;
        }
    }
    out_rowptr[n] = w;
    return w;
}

/* ------------------------------------------------------------------------ */
/* Counting.                                                                  */
/* ------------------------------------------------------------------------ */
typedef struct {
    u32 W;          /* words per bitmap row */
    u64 *rows;      /* R x W */
    u64 *stack;     /* (k + 2) levels x 2 x W scratch */
    u32 *list;      /* task vertex list (N+(u) or T) */
    u32 k;
    int ovf;        /* a needed binomial did not fit in 64 bits */
    int tails;      /* close SCT subtrees with held >= k-2 in closed form */
    u128 bytes;     /* algorithmic bytes of the processed tasks */
} ctx_t;

static inline int popc64(u64 x) { return __builtin_popcountll(x); }

static int popc_set(const u64 *a, u32 W)
{
    int c = 0;
    for (u32 w = 0; w < W; ++w) c += popc64(a[w]);
    return c;
}

/* Bitmap encode of the induced subgraph on list[0..R) (sorted ascending):
 * bit j of row i <=> list[j] in N+(list[i])  (warp_sorted_count_and_encode,
 * include/TriCountPrim.cuh:1768-1799); with full=1 also bit i of row j
 * (warp_sorted_count_and_encode_full, :1864-1900).  The reference searches
 * each list element in N+(list[i]); here the two sorted lists are merged. */
static void encode(ctx_t *c, const u64 *rowptr, const u32 *col, u32 R, int full)
{
    memset(c->rows, 0, sizeof(u64) * (size_t)R * c->W);
    for (u32 i = 0; i < R; ++i) {
        u32 w = c->list[i];
        const u32 *nw = col + rowptr[w];
        u64 dw = rowptr[w + 1] - rowptr[w];
        c->bytes += 8 + 4 * (u128)dw;
        u64 a = 0; u32 j = 0;
        while (a < dw && j < R) {
            if (nw[a] < c->list[j]) ++a;
            else if (nw[a] > c->list[j]) ++j;
            else {
                c->rows[(size_t)i * c->W + (j >> 6)] |= 1ull << (j & 63);
                if (full) c->rows[(size_t)j * c->W + (i >> 6)] |= 1ull << (i & 63);
                ++a; ++j;
            }
        }
    }
}

/* Orientation DFS (kclique/kckernels_orientation.cuh:381-463): the number of
 * `need`-vertex cliques of the DAG bitmap inside candidate set `cand`.
 * need==1 is the popcount that the reference adds at l+1==KCCOUNT; a child
 * is only entered when it still has `need-1` candidates (:449). */
static u128 dag_count(ctx_t *c, const u64 *cand, u32 need, u32 depth)
{
    u32 W = c->W;
    if (need == 1) return (u128)popc_set(cand, W);
    u128 total = 0;
    u64 *child = c->stack + (size_t)depth * 2 * W;
    for (u32 wi = 0; wi < W; ++wi) {
        u64 bits = cand[wi];
        while (bits) {
            u32 j = wi * 64 + (u32)__builtin_ctzll(bits);
            bits &= bits - 1;
            const u64 *rj = c->rows + (size_t)j * W;
            int cnt = 0;
            for (u32 w = 0; w < W; ++w) { child[w] = cand[w] & rj[w]; cnt += popc64(child[w]); }
            if ((u32)cnt >= need - 1) total += dag_count(c, child, need - 1, depth + 1);
        }
    }
    return total;
}

/* Pivoting (SCT) recursion (kclique/kckernels_pivoting.cuh:163-338):
 *  - pivot = smallest index among the candidates j that maximise |C & row_j|
 *    (:134-161 for the root, :250-309 below it);
 *  - branches = C \ row_pivot (the pivot is one of them, :169, :315);
 *  - a branch on the pivot adds a pivot vertex ("drop"), any other branch a
 *    held vertex (:206-209); held > k prunes (:214);
 *  - each child excludes the branch vertices taken before it (:237);
 *  - an empty child is a leaf that adds C(pivots, k - held) (:283-287,
 *    written there as nCR[drop][rsize - k]). */
static u128 sct(ctx_t *c, const u64 *C, u32 held, u32 piv, u32 depth)
{
    u32 W = c->W, k = c->k;
    if (held > k) return 0;
    int any = 0;
    for (u32 w = 0; w < W; ++w) if (C[w]) { any = 1; break; }
    if (!any) {
        u32 r = k - held;
        if (r > piv) return 0;
        int o = 0;
        u128 v = binom128(piv, r);
        if (v > (u128)UINT64_MAX) { o = 1; }
        if (o) c->ovf = 1;
        return v;
    }
    u32 cs = (u32)popc_set(C, W);
    /* closed-form tails (not in the reference): a subtree with h held and
     * p pivot vertices over candidates C holds sum_j C(p, k-h-j) c_j(C)
     * k-cliques, c_j = j-cliques of G[C]; c_0 = 1, c_1 = |C|, c_2 and c_3
     * below.  Tails close every node with h >= k-3. */
    if (c->tails && held == k) return 1;
    if (c->tails && held + 1 == k) return (u128)piv + cs;
    /* a clique below this node has at most held + piv + |C| vertices */
    if (held + piv + cs < k) return 0;
    /* pivot choice */
    int best = -1; u32 bj = 0;
    u128 score_sum = 0;
    for (u32 wi = 0; wi < W; ++wi) {
        u64 bits = C[wi];
        while (bits) {
            u32 j = wi * 64 + (u32)__builtin_ctzll(bits);
            bits &= bits - 1;
            const u64 *rj = c->rows + (size_t)j * W;
            int cnt = 0;
            for (u32 w = 0; w < W; ++w) cnt += popc64(C[w] & rj[w]);
            score_sum += (u128)cnt;
            if (cnt > best) { best = cnt; bj = j; }
        }
    }
    /* c_2 = |E(C)| = half the sum of the scores */
    if (c->tails && held + 2 == k)
        return binom128(piv, 2) + (u128)piv * cs + score_sum / 2;
    /* c_3 = triangles i < j < l of G[C], each counted at its smallest vertex */
    if (c->tails && held + 3 == k) {
        u128 tri = 0;
        for (u32 wi = 0; wi < W; ++wi) {
            u64 bi = C[wi];
            while (bi) {
                u32 i = wi * 64 + (u32)__builtin_ctzll(bi);
                bi &= bi - 1;
                const u64 *ri = c->rows + (size_t)i * W;
                for (u32 wj = 0; wj < W; ++wj) {
                    u64 bj2 = C[wj] & ri[wj];
                    if (wj < (i >> 6)) continue;
                    if (wj == (i >> 6)) bj2 &= (i & 63) == 63 ? 0 : (~0ull << ((i & 63) + 1));
                    while (bj2) {
                        u32 j = wj * 64 + (u32)__builtin_ctzll(bj2);
                        bj2 &= bj2 - 1;
                        const u64 *rj = c->rows + (size_t)j * W;
                        for (u32 w = j >> 6; w < W; ++w) {
                            u64 x = C[w] & ri[w] & rj[w];
                            if (w == (j >> 6)) x &= (j & 63) == 63 ? 0 : (~0ull << ((j & 63) + 1));
                            tri += (u128)popc64(x);
                        }
                    }
                }
            }
        }
        return binom128(piv, 3) + binom128(piv, 2) * cs + (u128)piv * (score_sum / 2) + tri;
    }
    /* ... and at most held + piv + 1 + max_j |C & row_j| (the largest
     * clique inside C has at most maxIntersection + 1 vertices; the edge
     * kernel's prune, kclique/kckernels.cuh:2573-2584) */
    if (held + piv + 1 + (u32)best < k) return 0;
    u64 *Cc = c->stack + (size_t)depth * 3 * W;      /* remaining candidates */
    u64 *B = Cc + W;                                 /* branch set */
    u64 *child = B + W;
    const u64 *rp = c->rows + (size_t)bj * W;
    for (u32 w = 0; w < W; ++w) { Cc[w] = C[w]; B[w] = C[w] & ~rp[w]; }
    u128 total = 0;
    for (u32 wi = 0; wi < W; ++wi) {
        while (B[wi]) {
            u32 b = wi * 64 + (u32)__builtin_ctzll(B[wi]);
            B[wi] &= B[wi] - 1;
            const u64 *rb = c->rows + (size_t)b * W;
            for (u32 w = 0; w < W; ++w) child[w] = Cc[w] & rb[w];
            if (b == bj) total += sct(c, child, held, piv + 1, depth + 1);
            else total += sct(c, child, held + 1, piv, depth + 1);
            Cc[wi] &= ~(1ull << (b & 63));
        }
    }
    return total;
}

/* Intersection of two sorted lists into out (block_count_and_set_tri,
 * include/TriCountPrim.cuh:2071-2103; ascending order here, the reference
 * appends in atomic order). */
static u32 intersect(const u32 *a, u64 la, const u32 *b, u64 lb, u32 *out)
{
    u64 i = 0, j = 0; u32 n = 0;
    while (i < la && j < lb) {
        if (a[i] < b[j]) ++i;
        else if (a[i] > b[j]) ++j;
        else { out[n++] = a[i]; ++i; ++j; }
    }
    return n;
}

/*
 * orc_count: number of k-cliques of the graph whose oriented CSR is
 * (rowptr, col), by one of the reference's four traversals.
 *   algo: 0 = orientation DFS, 1 = pivoting (SCT), 2 = pivoting with the
 *         closed-form tails used by the HIP kernels (same counts)
 *   gran: 0 = node tasks, 1 = edge tasks
 * Tasks are numbered 0..n-1 (node) or 0..m-1 (edge, in CSR order); only the
 * tasks with index % stride == offset are processed (stride=1: all), which
 * gives the bounded CPU-baseline sample.  Task filters: node tasks need
 * d+(u) >= k-1 (kclique/kclique.cuh:1610-1614); edge tasks need
 * |T| >= k-2 where T = N+(u) & N+(v) (the reference filters on
 * max(d+(u), d+(v)), kclique/kclique.cuh:2344-2355 with common.cuh:18-46,
 * which keeps tasks that cannot hold a clique).
 * nthreads <= 0 uses the OpenMP default.
 * Returns 0; *count gets the low 64 bits, *overflow is set when the count or
 * a binomial it needs is >= 2^64.  *tasks_done gets the number of tasks that
 * passed the filter; *alg_bytes the algorithmic bytes of those tasks
 * (8 + 4 d+(u) per own list, 8 + 4 d+(w) per neighbour list read).
 */
int orc_count(u32 n, const u64 *rowptr, const u32 *col, u32 k, int algo, int gran,
              u64 stride, u64 offset, int nthreads,
              u64 *count, int *overflow, u64 *tasks_done, u64 *alg_bytes)
{
    if (k < 1) return -1;
    if (stride == 0) stride = 1;
    u64 m = rowptr[n];
    u32 maxd = 0;
    for (u32 u = 0; u < n; ++u) {
        u64 d = rowptr[u + 1] - rowptr[u];
        if (d > maxd) maxd = (u32)d;
    }
    *overflow = 0;
    if (k == 1 || k == 2) {
        /* closed forms: n vertices, m edges (the sampled share of them) */
        u64 c = 0, t = 0;
        if (k == 1) { for (u64 u = offset; u < n; u += stride) { c++; t++; } }
        else if (gran == 0) { for (u64 u = offset; u < n; u += stride) { c += rowptr[u + 1] - rowptr[u]; t++; } }
        else { for (u64 e = offset; e < m; e += stride) { c++; t++; } }
        *count = c; *tasks_done = t; *alg_bytes = 0;
        return 0;
    }
    int nt = 1;
#ifdef _OPENMP
    nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#else
    (void)nthreads;
#endif
    u128 *part = (u128 *)calloc((size_t)nt, sizeof(u128));
    u128 *pbytes = (u128 *)calloc((size_t)nt, sizeof(u128));
    u64 *ptasks = (u64 *)calloc((size_t)nt, sizeof(u64));
    int *povf = (int *)calloc((size_t)nt, sizeof(int));
    u64 ntasks = (gran == 0) ? n : m;
    u32 *src_of = NULL;
    if (gran == 1) {
        src_of = (u32 *)malloc(sizeof(u32) * (m ? m : 1));
        for (u32 u = 0; u < n; ++u)
            for (u64 e = rowptr[u]; e < rowptr[u + 1]; ++e) src_of[e] = u;
    }
    u64 nsample = (ntasks > offset) ? (ntasks - offset + stride - 1) / stride : 0;

#ifdef _OPENMP
#pragma omp parallel num_threads(nt)
#endif
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        ctx_t c;
        c.k = k;
        c.W = 0;
        u32 Wmax = (maxd + 63) / 64;
        if (Wmax == 0) Wmax = 1;
        c.rows = (u64 *)malloc(sizeof(u64) * ((size_t)maxd + 1) * Wmax);
        c.stack = (u64 *)malloc(sizeof(u64) * ((size_t)maxd + k + 4) * 3 * Wmax);
        c.list = (u32 *)malloc(sizeof(u32) * ((size_t)maxd + 1));
        c.ovf = 0;
        c.tails = (algo == 2);
        c.bytes = 0;
        u128 acc = 0;
        u64 done = 0;
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 16)
#endif
        for (u64 s = 0; s < nsample; ++s) {
            u64 t = offset + s * stride;
            u32 R;
            if (gran == 0) {
                u32 u = (u32)t;
                u64 d = rowptr[u + 1] - rowptr[u];
                if (d < k - 1) continue;
                R = (u32)d;
                memcpy(c.list, col + rowptr[u], sizeof(u32) * R);
                c.bytes += 8 + 4 * (u128)d;
            } else {
                u32 u = src_of[t], v = col[t];
                u64 du = rowptr[u + 1] - rowptr[u], dv = rowptr[v + 1] - rowptr[v];
                if (du < k - 2 || dv < k - 2) continue;
                R = intersect(col + rowptr[u], du, col + rowptr[v], dv, c.list);
                if (R < k - 2) continue;
                c.bytes += 24 + 4 * (u128)(du + dv);
                /* k=3 orientation adds |T| without a bitmap
                 * (kckernels_orientation.cuh:546-547) */
                if (algo == 0 && k == 3) { done++; acc += R; continue; }
            }
            done++;
            c.W = (R + 63) / 64;
            if (c.W == 0) c.W = 1;
            encode(&c, rowptr, col, R, algo != 0);
            if (algo == 0) {
                u32 need = (gran == 0) ? k - 1 : k - 2;   /* vertices still to pick (>= 1) */
                if (gran == 0) {
                    /* second level: each row j is one subtree (:381-413) */
                    for (u32 j = 0; j < R; ++j) {
                        const u64 *rj = c.rows + (size_t)j * c.W;
                        if ((u32)popc_set(rj, c.W) >= need - 1)
                            acc += dag_count(&c, rj, need - 1, 0);
                    }
                } else {
                    /* deeper levels start from each row of T (:562-633) */
                    for (u32 j = 0; j < R; ++j) {
                        const u64 *rj = c.rows + (size_t)j * c.W;
                        if ((u32)popc_set(rj, c.W) >= need - 1)
                            acc += dag_count(&c, rj, need - 1, 0);
                    }
                }
            } else {
                /* root: held = 1 (the node) or 2 (the edge), C = all of list */
                u64 *C = c.stack + (size_t)(maxd + k + 3) * 3 * Wmax - Wmax;
                memset(C, 0, sizeof(u64) * c.W);
                for (u32 j = 0; j < R; ++j) C[j >> 6] |= 1ull << (j & 63);
                acc += sct(&c, C, gran == 0 ? 1 : 2, 0, 0);
            }
        }
        part[tid] = acc;
        pbytes[tid] = c.bytes;
        ptasks[tid] = done;
        povf[tid] = c.ovf;
        free(c.rows); free(c.stack); free(c.list);
    }
    u128 total = 0, tb = 0; u64 td = 0; int ov = 0;
    for (int i = 0; i < nt; ++i) { total += part[i]; tb += pbytes[i]; td += ptasks[i]; ov |= povf[i]; }
    if (total > (u128)UINT64_MAX) ov = 1;
    *count = (u64)total;
    *overflow = ov;
    *tasks_done = td;
    *alg_bytes = (u64)tb;
    free(part); free(pbytes); free(ptasks); free(povf); free(src_of);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Local (per-vertex) counts, `-m lkc` (kclique/kclique_local.cuh:24-639,    */
/* kernels kclique/kckernels_local.cuh): local[v] = number of k-cliques that  */
/* contain v.  The reference adds, per found clique or SCT leaf, to each     */
/* member: orientation adds 1 per member of an enumerated clique; pivoting   */
/* splits a leaf with h held and p pivot vertices into C(p, k-h) for each     */
/* held vertex and C(p-1, k-h-1) for each pivot vertex.                       */
/* ------------------------------------------------------------------------ */
typedef struct {
    ctx_t c;
    u128 *loc;        /* per-thread local counts, n entries */
    u32 roots[2];     /* the task's root vertices (held) */
    u32 nroots;
    u32 *H, nh;       /* held task-local indices on the current path */
    u32 *P, np;       /* pivot task-local indices on the current path */
} lctx_t;

static u128 binom_or0(lctx_t *L, int64_t n, int64_t r)
{
    if (r < 0 || n < 0 || r > n) return 0;
    u128 v = binom128((u32)n, (u32)r);
    if (v > (u128)UINT64_MAX) L->c.ovf = 1;
    return v;
}

static void add_held(lctx_t *L, u128 v)
{
    if (!v) return;
    for (u32 i = 0; i < L->nroots; ++i) L->loc[L->roots[i]] += v;
    for (u32 i = 0; i < L->nh; ++i) L->loc[L->c.list[L->H[i]]] += v;
}

/* Orientation DFS with the held path: `need` more vertices from cand. */
static void dag_local(lctx_t *L, const u64 *cand, u32 need, u32 depth)
{
    u32 W = L->c.W;
    if (need == 0) { add_held(L, 1); return; }
    if (need == 1) {
        u32 cnt = 0;
        for (u32 wi = 0; wi < W; ++wi) {
            u64 b = cand[wi];
            while (b) {
                u32 j = wi * 64 + (u32)__builtin_ctzll(b);
                b &= b - 1;
                L->loc[L->c.list[j]] += 1;
                cnt++;
            }
        }
        add_held(L, cnt);
        return;
    }
    u64 *child = L->c.stack + (size_t)depth * 2 * W;
    for (u32 wi = 0; wi < W; ++wi) {
        u64 bits = cand[wi];
        while (bits) {
            u32 j = wi * 64 + (u32)__builtin_ctzll(bits);
            bits &= bits - 1;
            const u64 *rj = L->c.rows + (size_t)j * W;
            int cnt = 0;
            for (u32 w = 0; w < W; ++w) { child[w] = cand[w] & rj[w]; cnt += popc64(child[w]); }
            if ((u32)cnt < need - 1) continue;
            L->H[L->nh++] = j;
            dag_local(L, child, need - 1, depth + 1);
            L->nh--;
        }
    }
}

/* SCT with the held and pivot sets of the path.  Without tails the leaves
 * carry the reference's hold/pivot split.  With tails a node with t = k - h
 * <= 3 closes its subtree: its cliques are H + S + Q with S a subset of the
 * pivots P and Q a clique of G[C], |S| + |Q| = t, so
 *   held vertex   sum_j C(p, t-j)   c_j(C)
 *   pivot vertex  sum_j C(p-1, t-1-j) c_j(C)
 *   x in C        sum_j C(p, t-j)   c_{j-1}(C & N(x))
 * with c_0 = 1, c_1 = size, c_2 = edges, c_3 = triangles. */
static void sct_local(lctx_t *L, const u64 *C, u32 depth)
{
    u32 W = L->c.W, k = L->c.k;
    u32 held = L->nroots + L->nh, piv = L->np;
    if (held > k) return;
    u32 cs = (u32)popc_set(C, W);
    int64_t t = (int64_t)k - held;
    if (cs == 0) {
        add_held(L, binom_or0(L, piv, t));
        u128 pv = binom_or0(L, (int64_t)piv - 1, t - 1);
        if (pv) for (u32 i = 0; i < L->np; ++i) L->loc[L->c.list[L->P[i]]] += pv;
        return;
    }
    if (held + piv + cs < k) return;
    if (L->c.tails && t <= 3) {
        /* per-candidate degree and edge count inside C */
        u128 c1 = cs, c2 = 0, c3 = 0;
        u64 *nx = L->c.stack + (size_t)depth * 3 * W;
        for (u32 wi = 0; wi < W; ++wi) {
            u64 bi = C[wi];
            while (bi) {
                u32 x = wi * 64 + (u32)__builtin_ctzll(bi);
                bi &= bi - 1;
                const u64 *rx = L->c.rows + (size_t)x * W;
                u32 dx = 0;
                for (u32 w = 0; w < W; ++w) { nx[w] = C[w] & rx[w]; dx += (u32)popc64(nx[w]); }
                u128 ex2 = 0;  /* twice the edges inside C & N(x) */
                if (t == 3)
                    for (u32 wj = 0; wj < W; ++wj) {
                        u64 bj = nx[wj];
                        while (bj) {
                            u32 y = wj * 64 + (u32)__builtin_ctzll(bj);
                            bj &= bj - 1;
                            const u64 *ry = L->c.rows + (size_t)y * W;
                            for (u32 w = 0; w < W; ++w) ex2 += (u128)popc64(nx[w] & ry[w]);
                        }
                    }
                c2 += dx;
                c3 += ex2 / 2;
                u128 add = binom_or0(L, piv, t - 1) + binom_or0(L, piv, t - 2) * dx + binom_or0(L, piv, t - 3) * (ex2 / 2);
                L->loc[L->c.list[x]] += add;
            }
        }
        c2 /= 2;
        c3 /= 3;
        u128 cj[4] = {1, c1, c2, c3};
        u128 tot = 0, ptot = 0;
        for (int64_t j = 0; j <= t; ++j) {
            tot += binom_or0(L, piv, t - j) * cj[j];
            ptot += binom_or0(L, (int64_t)piv - 1, t - 1 - j) * cj[j];
        }
        add_held(L, tot);
        if (ptot) for (u32 i = 0; i < L->np; ++i) L->loc[L->c.list[L->P[i]]] += ptot;
        return;
    }
    int best = -1; u32 bj = 0;
    for (u32 wi = 0; wi < W; ++wi) {
        u64 bits = C[wi];
        while (bits) {
            u32 j = wi * 64 + (u32)__builtin_ctzll(bits);
            bits &= bits - 1;
            const u64 *rj = L->c.rows + (size_t)j * W;
            int cnt = 0;
            for (u32 w = 0; w < W; ++w) cnt += popc64(C[w] & rj[w]);
            if (cnt > best) { best = cnt; bj = j; }
        }
    }
    if (held + piv + 1 + (u32)best < k) return;
    u64 *Cc = L->c.stack + (size_t)depth * 3 * W;
    u64 *B = Cc + W;
    u64 *child = B + W;
    const u64 *rp = L->c.rows + (size_t)bj * W;
    for (u32 w = 0; w < W; ++w) { Cc[w] = C[w]; B[w] = C[w] & ~rp[w]; }
    for (u32 wi = 0; wi < W; ++wi) {
        while (B[wi]) {
            u32 b = wi * 64 + (u32)__builtin_ctzll(B[wi]);
            B[wi] &= B[wi] - 1;
            const u64 *rb = L->c.rows + (size_t)b * W;
            for (u32 w = 0; w < W; ++w) child[w] = Cc[w] & rb[w];
            if (b == bj) { L->P[L->np++] = b; sct_local(L, child, depth + 1); L->np--; }
            else { L->H[L->nh++] = b; sct_local(L, child, depth + 1); L->nh--; }
            Cc[wi] &= ~(1ull << (b & 63));
        }
    }
}

/*
 * orc_count_local: local[v] = number of k-cliques of the oriented graph
 * (rowptr, col) that contain v, by the traversal `algo` (0 orientation,
 * 1 pivoting, 2 pivoting with closed-form tails) over node or edge tasks.
 * *overflow is set when an entry (or a binomial it needs) is >= 2^64.
 * Sum over v of local[v] = k * (the k-clique count).
 */
int orc_count_local(u32 n, const u64 *rowptr, const u32 *col, u32 k, int algo, int gran, int nthreads,
                    u64 *local, int *overflow)
{
    if (k < 1) return -1;
    u64 m = rowptr[n];
    *overflow = 0;
    if (k <= 2) {
        for (u32 v = 0; v < n; ++v) local[v] = (k == 1) ? 1 : rowptr[v + 1] - rowptr[v];
        if (k == 2) for (u64 e = 0; e < m; ++e) local[col[e]]++;
        return 0;
    }
    u32 maxd = 0;
    for (u32 u = 0; u < n; ++u) {
        u64 d = rowptr[u + 1] - rowptr[u];
        if (d > maxd) maxd = (u32)d;
    }
    int nt = 1;
#ifdef _OPENMP
    nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#else
    (void)nthreads;
#endif
    u128 *all = (u128 *)calloc((size_t)nt * (n ? n : 1), sizeof(u128));
    int *povf = (int *)calloc((size_t)nt, sizeof(int));
    u64 ntasks = (gran == 0) ? n : m;
    u32 *src_of = NULL;
    if (gran == 1) {
        src_of = (u32 *)malloc(sizeof(u32) * (m ? m : 1));
        for (u32 u = 0; u < n; ++u)
            for (u64 e = rowptr[u]; e < rowptr[u + 1]; ++e) src_of[e] = u;
    }
#ifdef _OPENMP
#pragma omp parallel num_threads(nt)
#endif
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        lctx_t L;
        memset(&L, 0, sizeof L);
        u32 Wmax = (maxd + 63) / 64;
        if (Wmax == 0) Wmax = 1;
        L.c.k = k;
        L.c.rows = (u64 *)malloc(sizeof(u64) * ((size_t)maxd + 1) * Wmax);
        L.c.stack = (u64 *)malloc(sizeof(u64) * ((size_t)maxd + k + 4) * 3 * Wmax);
        L.c.list = (u32 *)malloc(sizeof(u32) * ((size_t)maxd + 1));
        L.c.tails = (algo == 2);
        L.H = (u32 *)malloc(sizeof(u32) * ((size_t)maxd + k + 4));
        L.P = (u32 *)malloc(sizeof(u32) * ((size_t)maxd + k + 4));
        L.loc = all + (size_t)tid * n;
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 16)
#endif
        for (u64 t = 0; t < ntasks; ++t) {
            u32 R;
            if (gran == 0) {
                u32 u = (u32)t;
                u64 d = rowptr[u + 1] - rowptr[u];
                if (d < k - 1) continue;
                R = (u32)d;
                memcpy(L.c.list, col + rowptr[u], sizeof(u32) * R);
                L.roots[0] = u;
                L.nroots = 1;
            } else {
                u32 u = src_of[t], v = col[t];
                u64 du = rowptr[u + 1] - rowptr[u], dv = rowptr[v + 1] - rowptr[v];
                R = intersect(col + rowptr[u], du, col + rowptr[v], dv, L.c.list);
                if (R < k - 2) continue;
                L.roots[0] = u;
                L.roots[1] = v;
                L.nroots = 2;
            }
            L.c.W = (R + 63) / 64;
            if (L.c.W == 0) L.c.W = 1;
            encode(&L.c, rowptr, col, R, algo != 0);
            L.nh = L.np = 0;
            u64 *C = L.c.stack + (size_t)(maxd + k + 3) * 3 * Wmax - Wmax;
            memset(C, 0, sizeof(u64) * L.c.W);
            for (u32 j = 0; j < R; ++j) C[j >> 6] |= 1ull << (j & 63);
            if (algo == 0) dag_local(&L, C, k - L.nroots, 0);
            else sct_local(&L, C, 0);
        }
        povf[tid] = L.c.ovf;
        free(L.c.rows); free(L.c.stack); free(L.c.list); free(L.H); free(L.P);
    }
    int ov = 0;
    for (int i = 0; i < nt; ++i) ov |= povf[i];
    for (u32 v = 0; v < n; ++v) {
        u128 s = 0;
        for (int i = 0; i < nt; ++i) s += all[(size_t)i * n + v];
        if (s > (u128)UINT64_MAX) ov = 1;
        local[v] = (u64)s;
    }
    *overflow = ov;
    free(all); free(povf); free(src_of);
    return 0;
}

/*
 * orc_brute_force: k-cliques of a graph with n <= 64 vertices given as
 * adjacency bitmasks adj[v] (bit x set <=> edge v-x), by extending every
 * clique with larger vertex ids only.  This is the definition, not a
 * reference function.  *overflow is set above 2^64-1.
 */
static u128 bf_rec(const u64 *adj, u64 cand, u32 need)
{
    if (need == 0) return 1;
    u128 t = 0;
    while (cand) {
        u32 v = (u32)__builtin_ctzll(cand);
        cand &= cand - 1;
        t += bf_rec(adj, cand & adj[v], need - 1);
    }
    return t;
}

u64 orc_brute_force(u32 n, const u64 *adj, u32 k, int *overflow)
{
    u64 all = (n == 64) ? ~0ull : ((1ull << n) - 1);
    u128 t = (k == 0) ? 1 : bf_rec(adj, all, k);
    *overflow = t > (u128)UINT64_MAX;
    return (u64)t;
}

/* orc_threads: OpenMP threads the oracle would use. */
int orc_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
