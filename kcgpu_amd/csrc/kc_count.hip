// kc_count.hip -- exact k-clique counting kernels for gfx950 and kc_count().
//
// Reference kernels replaced (one wave of 64 lanes per task here, one
// 128-thread block per task there):
//   node orientation  kckernel_node_block_warp_binary_count   kclique/kckernels_orientation.cuh:319-475
//   edge orientation  kckernel_edge_block_warp_binary_count_o2 kclique/kckernels_orientation.cuh:479-650
//   node pivoting     kckernel_node_block_warp_binary_pivot_count kclique/kckernels_pivoting.cuh:2-346
//   edge pivoting     kckernel_edge_block_warp_binary_pivot_count kclique/kckernels.cuh:2263-2630
// plus the task selection of kclique/kclique.cuh:1318-1406 (degree-window
// buckets) -> size classes built by one device pass.
//
// Per task (a vertex u, or an oriented edge u->v) the wave
//   1. stages the task vertex list in LDS: N+(u), or T = N+(u) & N+(v)
//      (block_count_and_set_tri, include/TriCountPrim.cuh:2071-2103);
//   2. builds the induced subgraph as 64-bit bitmap rows in LDS: lanes stream
//      the concatenated out-lists of the task vertices (coalesced reads) and
//      binary-search each entry in the staged list (the reference searches the
//      other way, warp_sorted_count_and_encode, TriCountPrim.cuh:1768-1900);
//   3. counts:
//      orientation: P-lane partitions run independent DFS over the rows, a
//        child candidate set is the AND of two rows, the last level adds a
//        popcount (kckernels_orientation.cuh:381-463);
//      pivoting: one SCT tree per wave; the pivot is the smallest index among
//        the candidates j maximising |X & row_j|, scored by P-lane partitions;
//        leaves add C(pivots, k - held) from an exact uint64 table with
//        overflow bits.  Branches on non-pivot vertices are taken first and
//        the pivot branch last as a tail call, so the stack never holds more
//        than k - held + 2 levels (the reference keeps maxDeg levels in HBM).
// Sizes: tasks are binned by an upper bound of the list size (64, 128, 256,
// 512, 1024 rows in LDS; larger ones keep the same layout in an HBM slice).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "kc_device.h"

namespace kc {

constexpr int NCLASS = 6;
constexpr uint32_t CLASS_ROWS[NCLASS] = {64, 128, 256, 512, 1024, 0xFFFFFFFFu};
constexpr uint32_t INF32 = 0xFFFFFFFFu;
constexpr int ACC = 8;  // per-class accumulator slots

struct CountArgs {
    const uint32_t *rowptr;
    const uint32_t *col;
    const uint32_t *src;
    const uint32_t *tasks;
    uint32_t ntasks;
    uint32_t k;
    uint32_t psize;    // lanes per partition
    uint32_t nparts;   // partitions per wave used by the orientation DFS
    uint32_t chunk;    // tasks per queue grab
    uint32_t rmax;     // row capacity of the workspace
    uint32_t wmax;     // words per row capacity
    uint32_t levels;   // stack levels
    const uint64_t *binom;
    const uint8_t *binom_ovf;
    uint32_t bstride;  // binom[n * bstride + r]
    uint32_t *qhead;
    uint64_t *acc;     // [0] count [1] bytes [2] tasks [3] tree nodes [4] max task ticks [5] sum task ticks
    uint32_t *ovf;
    uint64_t *gscr;
    uint64_t gscr_words;
};

// ---- workspace layout (u64 units) -----------------------------------------
struct Layout {
    uint64_t misc, list, offs, starts, rows, stk, x, scal, total;
};

__host__ __device__ inline Layout make_layout(uint32_t rmax, uint32_t wmax, uint32_t levels, uint32_t nparts,
                                              bool pivot)
{
    Layout L;
    L.misc = 0;
    L.list = L.misc + 8;
    L.offs = L.list + (rmax + 1) / 2;
    L.starts = L.offs + (rmax + 2) / 2;
    L.rows = L.starts + (rmax + 1) / 2;
    L.stk = L.rows + (uint64_t)rmax * wmax;
    if (pivot) {
        L.x = L.stk + (uint64_t)levels * 2 * wmax;
        L.scal = L.x + wmax;
        L.total = L.scal + (uint64_t)levels * 2;
    } else {
        L.x = L.scal = L.stk + (uint64_t)nparts * levels * 2 * wmax;
        L.total = L.x;
    }
    return L;
}

// ---- wave helpers ------------------------------------------------------------
__device__ inline int lane_id() { return threadIdx.x & 63; }

__device__ inline uint64_t lanemask_lt()
{
    int l = lane_id();
    return l == 0 ? 0ull : (~0ull >> (64 - l));
}

__device__ inline uint32_t wave_incl_scan32(uint32_t x)
{
    int lane = lane_id();
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t y = __shfl_up(x, (unsigned)d, 64);
        if (lane >= d) x += y;
    }
    return x;
}

__device__ inline uint32_t wave_sum32(uint32_t x)
{
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) x += __shfl_xor(x, d, 64);
    return x;
}

__device__ inline uint32_t wave_min32(uint32_t x)
{
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) x = min(x, (uint32_t)__shfl_xor(x, d, 64));
    return x;
}

__device__ inline uint64_t wave_max64(uint64_t x)
{
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        uint64_t y = __shfl_xor(x, d, 64);
        x = y > x ? y : x;
    }
    return x;
}

// checked sum over the wave; *ovf set on carry-out
__device__ inline uint64_t wave_sum64_checked(uint64_t x, uint32_t &ovf)
{
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        uint64_t y = __shfl_xor(x, d, 64);
        uint64_t s = x + y;
        if (s < x) ovf = 1;
        x = s;
    }
    return x;
}

// reductions inside a partition of P lanes (P a power of two, aligned)
__device__ inline uint32_t part_min32(uint32_t x, uint32_t P)
{
    for (uint32_t d = P >> 1; d >= 1; d >>= 1) x = min(x, (uint32_t)__shfl_xor(x, (int)d, 64));
    return x;
}

__device__ inline uint32_t part_sum32(uint32_t x, uint32_t P)
{
    for (uint32_t d = P >> 1; d >= 1; d >>= 1) x += __shfl_xor(x, (int)d, 64);
    return x;
}

__device__ inline void add_checked(uint64_t &acc, uint64_t v, uint32_t &ovf)
{
    uint64_t s = acc + v;
    if (s < acc) ovf = 1;
    acc = s;
}

// first index in sorted a[0..n) with a[i] >= x
template <class P>
__device__ inline uint32_t lower_bound_u32(const P *a, uint32_t n, uint32_t x)
{
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        uint32_t mid = (lo + hi) >> 1;
        if (a[mid] < x) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// ---- step 2: bitmap rows of the induced subgraph on list[0..R) -------------
// bit j of row i <=> list[j] in N+(list[i]); FULL also sets bit i of row j.
template <bool FULL>
__device__ void build_rows(const CountArgs &a, uint32_t *list, uint32_t *offs, uint32_t *starts, uint64_t *rows,
                           uint32_t R, uint32_t W, uint64_t &bytes)
{
    const int lane = lane_id();
    for (uint32_t i = lane; i < R * W; i += 64) rows[i] = 0;
    uint32_t carry = 0;
    for (uint32_t base = 0; base < R; base += 64) {
        uint32_t i = base + lane, d = 0;
        if (i < R) {
            uint32_t w = list[i];
            uint32_t s = a.rowptr[w];
            d = a.rowptr[w + 1] - s;
            starts[i] = s;
        }
        uint32_t inc = wave_incl_scan32(d);
        if (i < R) offs[i] = carry + inc - d;
        carry += __shfl(inc, 63, 64);
    }
    if (lane == 0) offs[R] = carry;
    __syncthreads();
    const uint32_t total = carry;
    bytes += 8ull * R + 4ull * total;
    const uint32_t lo_v = list[0], hi_v = list[R - 1];
    for (uint32_t f = lane; f < total; f += 64) {
        // row i: the last i with offs[i] <= f
        uint32_t lo = 0, hi = R - 1;
        while (lo < hi) {
            uint32_t mid = (lo + hi + 1) >> 1;
            if (offs[mid] <= f) lo = mid;
            else hi = mid - 1;
        }
        uint32_t x = a.col[starts[lo] + (f - offs[lo])];
        if (x < lo_v || x > hi_v) continue;
        uint32_t j = lower_bound_u32(list, R, x);
        if (j < R && list[j] == x) {
            atomicOr((unsigned long long *)&rows[(uint64_t)lo * W + (j >> 6)], 1ull << (j & 63));
            if (FULL) atomicOr((unsigned long long *)&rows[(uint64_t)j * W + (lo >> 6)], 1ull << (lo & 63));
        }
    }
    __syncthreads();
}

// ---- step 3a: orientation DFS ---------------------------------------------
// Number of r-vertex cliques of the DAG rows (r >= 2).  Partitions of P lanes
// take first-level rows from a per-task LDS counter (*next_row) and run an
// independent DFS each; a lane only ever touches its own words (w = lip,
// lip + P, ...), so the DFS needs no barrier.  Returns the lane's share.
__device__ uint64_t orient_count(const CountArgs &a, const uint64_t *rows, uint64_t *stk, uint32_t *next_row,
                                 uint32_t R, uint32_t W, uint32_t r, uint64_t &nodes, uint32_t &ovf)
{
    const int lane = lane_id();
    uint64_t cnt = 0;
    if (r == 2) {
        for (uint32_t i = lane; i < R * W; i += 64) cnt += (uint64_t)__popcll(rows[i]);
        return cnt;
    }
    const uint32_t P = a.psize, NP = a.nparts;
    const uint32_t part = lane / P, lip = lane % P;
    if (part >= NP) return 0;
    const uint32_t wmax = a.wmax;
    uint64_t *base = stk + (uint64_t)part * a.levels * 2 * wmax;
    const uint32_t rem0 = r - 1;  // vertices still to pick below a first-level row
    uint64_t my_nodes = 0;
    for (;;) {
        uint32_t j = 0;
        if (lip == 0) j = atomicAdd(next_row, 1u);
        j = __shfl(j, (int)(part * P), 64);
        if (j >= R) break;
        // level 0: C = I = row j
        uint32_t c0 = 0;
        for (uint32_t w = lip; w < W; w += P) {
            uint64_t v = rows[(uint64_t)j * W + w];
            base[w] = v;
            base[wmax + w] = v;
            c0 += (uint32_t)__popcll(v);
        }
        c0 = part_sum32(c0, P);
        if (c0 < rem0) continue;
        int t = 0;
        for (;;) {
            uint64_t *C = base + (uint64_t)t * 2 * wmax, *I = C + wmax;
            uint32_t b = INF32;
            for (uint32_t w = lip; w < W; w += P) {
                uint64_t x = I[w];
                if (x) {
                    b = w * 64 + (uint32_t)__builtin_ctzll(x);
                    break;
                }
            }
            b = part_min32(b, P);
            if (b == INF32) {
                if (t == 0) break;
                --t;
                continue;
            }
            const uint32_t bw = b >> 6;
            if (bw % P == lip) I[bw] &= ~(1ull << (b & 63));
            uint64_t *Cn = C + 2 * wmax;
            const uint64_t *rb = rows + (uint64_t)b * W;
            uint32_t c = 0;
            for (uint32_t w = lip; w < W; w += P) {
                uint64_t x = C[w] & rb[w];
                Cn[w] = x;
                Cn[wmax + w] = x;
                c += (uint32_t)__popcll(x);
            }
            c = part_sum32(c, P);
            ++my_nodes;
            const uint32_t remc = rem0 - 1 - (uint32_t)t;  // picks left below b
            if (remc == 1) {
                if (lip == 0) add_checked(cnt, c, ovf);
            } else if (c >= remc) {
                ++t;
            }
        }
    }
    if (lip == 0) nodes += my_nodes;
    return cnt;
}

// ---- step 3b: pivoting (SCT) ----------------------------------------------
struct Scal {
    uint32_t held, piv, pivot, pad;
};

// A subtree rooted at a node with h held vertices, p pivot vertices and
// candidate set X represents every clique R_h + S + Q with S a clique of G[X]
// and Q a subset of the pivots, so its k-cliques number
//     sum_j C(p, k-h-j) * c_j(X),   c_j = j-cliques of G[X].
// With c_0 = 1, c_1 = |X| and c_2 = |E(X)| = (1/2) sum_{j in X} |X & row_j|
// (the pivot scores) the subtree is closed-form once h >= k-2:
//     h = k:   1        h = k-1:   p + |X|        h = k-2:  C(p,2) + p|X| + |E(X)|
// so the SCT is only expanded while h < k-2.  The register-bitmap kernel also
// closes h = k-3 with c_3 = triangles of G[X].  A node is also dropped when
// h + p + max_j|X & row_j| + 1 < k (the largest clique inside X has at most
// maxIntersection + 1 vertices, cf. kclique/kckernels.cuh:2573-2584).
__device__ uint64_t pivot_count(const CountArgs &a, const uint64_t *rows, uint64_t *stk, uint64_t *X, Scal *scal,
                                uint32_t *cand, uint32_t R, uint32_t W, uint32_t h0, uint64_t &nodes, uint32_t &ovf)
{
    const int lane = lane_id();
    const uint64_t lt = lanemask_lt();
    const uint32_t k = a.k, wmax = a.wmax;
    const uint32_t P = a.psize, NPp = 64 / P, part = lane / P, lip = lane % P;
    uint64_t acc = 0;
    // root candidate set = all R list vertices
    for (uint32_t w = lane; w < W; w += 64) {
        uint32_t lo = w * 64;
        X[w] = (R - lo >= 64) ? ~0ull : ((1ull << (R - lo)) - 1);
    }
    __syncthreads();
    int t = -1;
    bool pending = true;
    uint32_t h = h0, p = 0;
    uint64_t my_nodes = 0;
    for (;;) {
        if (pending) {
            pending = false;
            ++my_nodes;
            // candidate list of X (uniform word loop: the ballot of a word is the word)
            uint32_t cx = 0;
            for (uint32_t w = 0; w < W; ++w) {
                const uint64_t xw = X[w];
                if ((xw >> lane) & 1ull) cand[cx + (uint32_t)__popcll(xw & lt)] = w * 64 + (uint32_t)lane;
                cx += (uint32_t)__popcll(xw);
            }
            __syncthreads();
            if (cx == 0) {
                if (k >= h && k - h <= p) {
                    uint64_t idx = (uint64_t)p * a.bstride + (k - h);
                    if (a.binom_ovf[idx]) ovf = 1;
                    add_checked(acc, a.binom[idx], ovf);
                }
            } else if (h >= k) {
                add_checked(acc, 1, ovf);
            } else if (h + 1 == k) {
                add_checked(acc, (uint64_t)p + cx, ovf);
            } else if (h + p + cx >= k) {
                // scores |X & row_j| for j in X: P lanes per candidate
                uint32_t best_s = 0, best_j = INF32;
                uint64_t sum_s = 0;
                for (uint32_t c = part; c < cx; c += NPp) {
                    const uint32_t j = cand[c];
                    const uint64_t *rj = rows + (uint64_t)j * W;
                    uint32_t sc = 0;
                    for (uint32_t w = lip; w < W; w += P) sc += (uint32_t)__popcll(X[w] & rj[w]);
                    sc = part_sum32(sc, P);
                    sum_s += sc;
                    if (best_j == INF32 || sc > best_s) {
                        best_s = sc;
                        best_j = j;
                    }
                }
                uint64_t key = (lip == 0 && best_j != INF32) ? (((uint64_t)best_s << 32) | (INF32 - best_j)) : 0ull;
                key = wave_max64(key);
                uint32_t dummy = 0;
                const uint64_t tot_s = wave_sum64_checked(lip == 0 ? sum_s : 0ull, dummy);
                const uint32_t max_s = (uint32_t)(key >> 32);
                if (h + 2 == k) {
                    add_checked(acc, (uint64_t)p * (p - (p ? 1 : 0)) / 2, ovf);
                    add_checked(acc, (uint64_t)p * cx, ovf);
                    add_checked(acc, tot_s / 2, ovf);
                } else if (h + p + max_s + 1 >= k) {
                    const uint32_t pv = INF32 - (uint32_t)(key & 0xFFFFFFFFull);
                    ++t;
                    uint64_t *Cc = stk + (uint64_t)t * 2 * wmax, *B = Cc + wmax;
                    const uint64_t *rp = rows + (uint64_t)pv * W;
                    for (uint32_t w = lane; w < W; w += 64) {
                        uint64_t x = X[w];
                        Cc[w] = x;
                        B[w] = x & ~rp[w];
                    }
                    if (lane == 0) scal[t] = Scal{h, p, pv, 0};
                    __syncthreads();
                }
            }
        }
        if (t < 0) break;
        // next branch of level t: non-pivot vertices first, the pivot last
        uint64_t *Cc = stk + (uint64_t)t * 2 * wmax, *B = Cc + wmax;
        const Scal sc = scal[t];
        uint32_t b = INF32;
        for (uint32_t w = lane; w < W; w += 64) {
            uint64_t x = B[w];
            if (w == (sc.pivot >> 6)) x &= ~(1ull << (sc.pivot & 63));
            if (x) {
                b = w * 64 + (uint32_t)__builtin_ctzll(x);
                break;
            }
        }
        b = wave_min32(b);
        const bool last = (b == INF32);
        if (last) b = sc.pivot;
        const uint64_t *rb = rows + (uint64_t)b * W;
        for (uint32_t w = lane; w < W; w += 64) {
            uint64_t c = Cc[w];
            X[w] = c & rb[w];
            if (w == (b >> 6)) {
                Cc[w] = c & ~(1ull << (b & 63));
                B[w] &= ~(1ull << (b & 63));
            }
        }
        __syncthreads();
        h = sc.held + (b != sc.pivot ? 1u : 0u);
        p = sc.piv + (b == sc.pivot ? 1u : 0u);
        if (last) --t;  // the pivot branch replaces its level
        if (h <= k) pending = true;
    }
    if (lane == 0) nodes += my_nodes;
    return acc;
}

// ---- step 3c: pivoting with the bitmap in registers (R <= 64*S) -------------
// Lane l holds rows l, l+64, ..., l+64(S-1) of the full bitmap in VGPRs, so a
// pivot score is S*S AND+popcount on the lane's own registers; the rows a
// branch needs whole are broadcast-read from the LDS copy left by the build.  The SCT
// state is uniform: candidate set X and the level variables live in scalar
// registers, and the stack is spread across the lanes (lane t keeps level t in
// VGPRs, read back with v_readlane).  Arg-max and sums use ballots: the max
// by a bitwise search, the sum bit-sliced, so no LDS and no shuffles are
// touched after the rows are loaded.
__device__ inline uint64_t rdlane64(uint64_t v, uint32_t src)
{
    uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, (int)src);
    uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), (int)src);
    return ((uint64_t)hi << 32) | lo;
}

__device__ inline uint32_t ballot_sum(uint32_t v, int bits)
{
    uint32_t s = 0;
    for (int b = 0; b < bits; ++b) s += (uint32_t)__popcll(__ballot((v >> b) & 1u)) << b;
    return s;
}

// word w of row b from the LDS copy (uniform address: a broadcast read)
__device__ inline uint64_t lds_row_word(const uint64_t *rows, uint32_t W, uint32_t b, int w)
{
    if ((uint32_t)w >= W) return 0ull;
    const uint64_t v = rows[(uint64_t)b * W + w];
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}

template <int S>
__device__ uint64_t pivot_reg(const CountArgs &a, const uint64_t *rows, uint32_t R, uint32_t h0, uint64_t &nodes,
                              uint32_t &ovf)
{
    constexpr int NB = (S == 1) ? 7 : (S == 2) ? 8 : 9;  // bits of (score + 1) <= 64*S
    const uint32_t lane = (uint32_t)lane_id();
    const uint32_t k = a.k;
    const uint32_t W = (R + 63) / 64;
    uint64_t row[S][S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const uint32_t j = lane + 64u * s;
#pragma unroll
        for (int w = 0; w < S; ++w) row[s][w] = (j < R && (uint32_t)w < W) ? rows[(uint64_t)j * W + w] : 0ull;
    }
    uint64_t X[S], stkC[S], stkB[S];
    uint32_t stkM = 0;
#pragma unroll
    for (int w = 0; w < S; ++w) {
        const uint32_t lo = 64u * w;
        X[w] = R <= lo ? 0ull : (R - lo >= 64 ? ~0ull : ((1ull << (R - lo)) - 1));
        stkC[w] = stkB[w] = 0;
    }
    int t = -1;
    bool pending = true;
    uint32_t h = h0, p = 0;
    uint64_t acc = 0, my_nodes = 0;
    for (;;) {
        if (pending) {
            pending = false;
            ++my_nodes;
            uint32_t cx = 0;
#pragma unroll
            for (int w = 0; w < S; ++w) cx += (uint32_t)__popcll(X[w]);
            if (cx == 0) {
                if (k >= h && k - h <= p) {
                    const uint64_t idx = (uint64_t)p * a.bstride + (k - h);
                    if (a.binom_ovf[idx]) ovf = 1;
                    add_checked(acc, a.binom[idx], ovf);
                }
            } else if (h >= k) {
                add_checked(acc, 1, ovf);
            } else if (h + 1 == k) {
                add_checked(acc, (uint64_t)p + cx, ovf);
            } else if (h + p + cx >= k) {
                uint32_t sv[S], lsum = 0;
#pragma unroll
                for (int s = 0; s < S; ++s) {
                    const bool mem = (X[s] >> lane) & 1ull;
                    uint32_t v = 0;
#pragma unroll
                    for (int w = 0; w < S; ++w) v += (uint32_t)__popcll(X[w] & row[s][w]);
                    sv[s] = mem ? v + 1 : 0;
                    lsum += mem ? v : 0;
                }
                uint32_t best = 0;
#pragma unroll
                for (int bt = NB; bt >= 0; --bt) {
                    const uint32_t c = best | (1u << bt);
                    bool any = false;
#pragma unroll
                    for (int s = 0; s < S; ++s) any |= sv[s] >= c;
                    if (__ballot(any)) best = c;
                }
                if (h + 3 >= k) {
                    const uint32_t e2 = ballot_sum(lsum, NB + 3);  // twice the edges of G[X]
                    const uint64_t pp = p;
                    if (h + 2 == k) {
                        add_checked(acc, pp * (pp - (pp ? 1 : 0)) / 2, ovf);
                        add_checked(acc, pp * cx, ovf);
                        add_checked(acc, e2 / 2, ovf);
                    } else {
                        // h = k-3: C(p,3) + C(p,2)|X| + p|E(X)| + triangles(X); each
                        // triangle i < j < l is counted once, at its smallest vertex
                        uint32_t tri = 0;
                        for (int wi = 0; wi < S; ++wi) {
                            uint64_t bits = X[wi];
                            while (bits) {
                                const uint32_t i = 64u * wi + (uint32_t)__builtin_ctzll(bits);
                                bits &= bits - 1;
                                uint64_t Y[S];
                                bool any = false;
#pragma unroll
                                for (int w = 0; w < S; ++w) {
                                    uint64_t above = (w < wi) ? 0ull : (w > wi) ? ~0ull
                                                   : ((i & 63) == 63 ? 0ull : (~0ull << ((i & 63) + 1)));
                                    Y[w] = X[w] & lds_row_word(rows, W, i, w) & above;
                                    any |= Y[w] != 0;
                                }
                                if (!any) continue;
#pragma unroll
                                for (int sl = 0; sl < S; ++sl) {
                                    if (!((Y[sl] >> lane) & 1ull)) continue;
                                    uint32_t v = 0;
#pragma unroll
                                    for (int w = 0; w < S; ++w) {
                                        uint64_t above = (w < sl) ? 0ull : (w > sl) ? ~0ull
                                                       : (lane == 63 ? 0ull : (~0ull << (lane + 1)));
                                        v += (uint32_t)__popcll(Y[w] & row[sl][w] & above);
                                    }
                                    tri += v;
                                }
                            }
                        }
                        tri = wave_sum32(tri);
                        add_checked(acc, pp * (pp - (pp ? 1 : 0)) * (pp >= 2 ? pp - 2 : 0) / 6, ovf);
                        add_checked(acc, pp * (pp - (pp ? 1 : 0)) / 2 * cx, ovf);
                        add_checked(acc, pp * (e2 / 2), ovf);
                        add_checked(acc, tri, ovf);
                    }
                } else if (h + p + best >= k) {  // best = max score + 1
                    uint32_t pv = 0;
#pragma unroll
                    for (int s = S - 1; s >= 0; --s) {
                        const uint64_t bal = __ballot(sv[s] == best);
                        if (bal) pv = 64u * s + (uint32_t)__builtin_ctzll(bal);
                    }
                    ++t;
#pragma unroll
                    for (int w = 0; w < S; ++w) {
                        const uint64_t bw = X[w] & ~lds_row_word(rows, W, pv, w);
                        stkC[w] = (lane == (uint32_t)t) ? X[w] : stkC[w];
                        stkB[w] = (lane == (uint32_t)t) ? bw : stkB[w];
                    }
                    const uint32_t M = h | (p << 8) | (pv << 20);
                    stkM = (lane == (uint32_t)t) ? M : stkM;
                }
            }
        }
        if (t < 0) break;
        // next branch of level t: non-pivot vertices first, the pivot last
        uint64_t Cc[S], B[S];
#pragma unroll
        for (int w = 0; w < S; ++w) {
            Cc[w] = rdlane64(stkC[w], (uint32_t)t);
            B[w] = rdlane64(stkB[w], (uint32_t)t);
        }
        const uint32_t M = (uint32_t)__builtin_amdgcn_readlane((int)stkM, t);
        const uint32_t hh = M & 0xFF, pp = (M >> 8) & 0xFFF, pv = M >> 20;
        uint32_t b = INF32;
#pragma unroll
        for (int w = S - 1; w >= 0; --w) {
            uint64_t x = B[w];
            if ((uint32_t)w == (pv >> 6)) x &= ~(1ull << (pv & 63));
            if (x) b = 64u * w + (uint32_t)__builtin_ctzll(x);
        }
        const bool last = (b == INF32);
        if (last) b = pv;
        const uint32_t bs = b >> 6, bl = b & 63;
#pragma unroll
        for (int w = 0; w < S; ++w) X[w] = Cc[w] & lds_row_word(rows, W, b, w);
        h = hh + (b != pv ? 1u : 0u);
        p = pp + (b == pv ? 1u : 0u);
        if (last) {
            --t;  // the pivot branch replaces its level
        } else {
#pragma unroll
            for (int w = 0; w < S; ++w) {
                if ((uint32_t)w == bs) {
                    Cc[w] &= ~(1ull << bl);
                    B[w] &= ~(1ull << bl);
                }
                stkC[w] = (lane == (uint32_t)t) ? Cc[w] : stkC[w];
                stkB[w] = (lane == (uint32_t)t) ? B[w] : stkB[w];
            }
        }
        if (h <= k) pending = true;
    }
    if (lane == 0) nodes += my_nodes;
    return acc;
}

// ---- step 1: stage a task's vertex list in list[] ------------------------------
// node task: N+(u); edge task: T = N+(u) & N+(v), walking the shorter list and
// binary-searching the longer one, compacted with a ballot (ascending order).
template <bool EDGE>
__device__ inline uint32_t stage_task(const CountArgs &a, uint32_t task, uint32_t *list, uint64_t &bytes)
{
    const int lane = lane_id();
    if (!EDGE) {
        const uint32_t b = a.rowptr[task], e = a.rowptr[task + 1];
        const uint32_t R = e - b;
        bytes += 8 + 4ull * R;
        for (uint32_t i = lane; i < R; i += 64) list[i] = a.col[b + i];
        return R;
    }
    const uint32_t u = a.src[task], v = a.col[task];
    const uint32_t ub = a.rowptr[u], ue = a.rowptr[u + 1], vb = a.rowptr[v], ve = a.rowptr[v + 1];
    const uint32_t du = ue - ub, dv = ve - vb;
    bytes += 24 + 4ull * (du + dv);
    const uint32_t *sa = a.col + (du <= dv ? ub : vb), *sb = a.col + (du <= dv ? vb : ub);
    const uint32_t la = min(du, dv), lb = max(du, dv);
    uint32_t cnt = 0;
    for (uint32_t i0 = 0; i0 < la; i0 += 64) {
        const uint32_t i = i0 + lane;
        bool hit = false;
        uint32_t x = 0;
        if (i < la) {
            x = sa[i];
            const uint32_t j = lower_bound_u32(sb, lb, x);
            hit = j < lb && sb[j] == x;
        }
        const uint64_t bal = __ballot(hit);
        if (hit) list[cnt + (uint32_t)__popcll(bal & lanemask_lt())] = x;
        cnt += (uint32_t)__popcll(bal);
    }
    return cnt;
}

// register-bitmap pivoting kernel for tasks of at most 64*S vertices
template <bool EDGE, int S>
__global__ __launch_bounds__(64) void k_pivot_reg(CountArgs a)
{
    extern __shared__ uint64_t lds[];
    const int lane = lane_id();
    const Layout L = make_layout(a.rmax, a.wmax, 0, 0, true);
    uint32_t *list = reinterpret_cast<uint32_t *>(lds + L.list);
    uint32_t *offs = reinterpret_cast<uint32_t *>(lds + L.offs);
    uint32_t *starts = reinterpret_cast<uint32_t *>(lds + L.starts);
    uint64_t *rows = lds + L.rows;
    uint64_t total = 0, bytes = 0, ntask = 0, nodes = 0, tmax = 0, tsum = 0;
    uint32_t ovf = 0;
    for (;;) {
        uint32_t start = 0;
        if (lane == 0) start = atomicAdd(a.qhead, a.chunk);
        start = __shfl(start, 0, 64);
        if (start >= a.ntasks) break;
        const uint32_t stop = min(a.ntasks, start + a.chunk);
        for (uint32_t ti = start; ti < stop; ++ti) {
            const uint64_t t0 = wall_clock64();
            ++ntask;
            const uint32_t R = stage_task<EDGE>(a, a.tasks[ti], list, bytes);
            __syncthreads();
            build_rows<true>(a, list, offs, starts, rows, R, (R + 63) / 64, bytes);
            add_checked(total, pivot_reg<S>(a, rows, R, EDGE ? 2u : 1u, nodes, ovf), ovf);
            __syncthreads();
            const uint64_t dt = wall_clock64() - t0;
            tmax = dt > tmax ? dt : tmax;
            tsum += dt;
        }
    }
    uint32_t novf = wave_min32(ovf ? 0u : 1u) == 0 ? 1u : 0u;
    uint32_t nd_wrap = 0;
    const uint64_t nd = wave_sum64_checked(nodes, nd_wrap);
    if (lane == 0) {
        unsigned long long old = atomicAdd((unsigned long long *)&a.acc[0], (unsigned long long)total);
        if (old + total < old) novf = 1;
        atomicAdd((unsigned long long *)&a.acc[1], (unsigned long long)bytes);
        atomicAdd((unsigned long long *)&a.acc[2], (unsigned long long)ntask);
        atomicAdd((unsigned long long *)&a.acc[3], (unsigned long long)nd);
        atomicMax((unsigned long long *)&a.acc[4], (unsigned long long)tmax);
        atomicAdd((unsigned long long *)&a.acc[5], (unsigned long long)tsum);
        if (novf) atomicOr(a.ovf, 1u);
    }
}

// ---- the kernel ---------------------------------------------------------------
template <bool PIVOT, bool EDGE, bool GLOBAL>
__global__ __launch_bounds__(64) void k_count(CountArgs a)
{
    extern __shared__ uint64_t lds[];
    const int lane = lane_id();
    uint64_t *ws = GLOBAL ? a.gscr + (uint64_t)blockIdx.x * a.gscr_words : lds;
    const Layout L = make_layout(a.rmax, a.wmax, a.levels, a.nparts, PIVOT);
    uint32_t *misc = reinterpret_cast<uint32_t *>(ws + L.misc);
    uint32_t *list = reinterpret_cast<uint32_t *>(ws + L.list);
    uint32_t *offs = reinterpret_cast<uint32_t *>(ws + L.offs);
    uint32_t *starts = reinterpret_cast<uint32_t *>(ws + L.starts);
    uint64_t *rows = ws + L.rows;
    uint64_t *stk = ws + L.stk;
    uint64_t *X = ws + L.x;
    Scal *scal = reinterpret_cast<Scal *>(ws + L.scal);

    const uint32_t k = a.k;
    uint64_t total = 0, bytes = 0, ntask = 0, nodes = 0;
    uint32_t ovf = 0;
    for (;;) {
        uint32_t start = 0;
        if (lane == 0) start = atomicAdd(a.qhead, a.chunk);
        start = __shfl(start, 0, 64);
        if (start >= a.ntasks) break;
        const uint32_t stop = min(a.ntasks, start + a.chunk);
        for (uint32_t ti = start; ti < stop; ++ti) {
            ++ntask;
            const uint32_t R = stage_task<EDGE>(a, a.tasks[ti], list, bytes);
            if (EDGE && !PIVOT && k == 3) {  // edge orientation at k=3 adds |T|
                add_checked(total, R, ovf);
                continue;
            }
            if (lane == 0) misc[0] = 0;
            __syncthreads();
            const uint32_t W = (R + 63) / 64;
            if (!PIVOT) {
                // (k-1)-cliques of the DAG on N+(u) / (k-2)-cliques of the DAG on T
                build_rows<false>(a, list, offs, starts, rows, R, W, bytes);
                uint64_t part = orient_count(a, rows, stk, misc, R, W, EDGE ? k - 2 : k - 1, nodes, ovf);
                uint64_t s = wave_sum64_checked(part, ovf);
                add_checked(total, s, ovf);
            } else {
                // SCT rooted at the node (held 1) or the edge (held 2)
                build_rows<true>(a, list, offs, starts, rows, R, W, bytes);
                uint64_t s = pivot_count(a, rows, stk, X, scal, offs, R, W, EDGE ? 2u : 1u, nodes, ovf);
                add_checked(total, s, ovf);
            }
            __syncthreads();
        }
    }
    // all lanes hold the same `total`; one atomic per wave
    uint32_t novf = wave_min32(ovf ? 0u : 1u) == 0 ? 1u : 0u;
    uint32_t nd_wrap = 0;
    const uint64_t nd = wave_sum64_checked(nodes, nd_wrap);  // tree nodes are per-lane shares
    if (lane == 0) {
        unsigned long long old = atomicAdd((unsigned long long *)&a.acc[0], (unsigned long long)total);
        if (old + total < old) novf = 1;
        atomicAdd((unsigned long long *)&a.acc[1], (unsigned long long)bytes);
        atomicAdd((unsigned long long *)&a.acc[2], (unsigned long long)ntask);
        atomicAdd((unsigned long long *)&a.acc[3], (unsigned long long)nd);
        if (novf) atomicOr(a.ovf, 1u);
    }
}

// |T| = |N+(u) & N+(v)| for every oriented edge (one lane per edge, merge)
__global__ __launch_bounds__(256) void k_tsize(const uint32_t *rowptr, const uint32_t *col, const uint32_t *src,
                                                uint32_t m, uint32_t *tsz)
{
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < m; e += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t u = src[e], v = col[e];
        uint32_t i = rowptr[u], ie = rowptr[u + 1], j = rowptr[v], je = rowptr[v + 1];
        uint32_t c = 0;
        while (i < ie && j < je) {
            const uint32_t x = col[i], y = col[j];
            c += (x == y);
            i += (x <= y);
            j += (y <= x);
        }
        tsz[e] = c;
    }
}

// ---- task classification ----------------------------------------------------
__device__ inline int class_of(uint32_t s)
{
    return s <= 64 ? 0 : s <= 128 ? 1 : s <= 256 ? 2 : s <= 512 ? 3 : s <= 1024 ? 4 : 5;
}

// size of a task: d+(u) for a node, |T| (from k_tsize) for an edge; tasks
// below `need` cannot hold a k-clique and are dropped
template <bool EDGE>
__global__ __launch_bounds__(256) void k_classify(const uint32_t *rowptr, const uint32_t *tsz,
                                                   uint32_t ntask, uint32_t need, uint32_t *cls_cnt, uint32_t *cls_max,
                                                   uint32_t *cls_pos, uint32_t *out, const uint32_t *cls_off,
                                                   int scatter)
{
    const int lane = threadIdx.x & 63;
    for (uint64_t base = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) & ~63ull; base < ntask;
         base += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t t = base + lane;
        int c = -1;
        uint32_t s = 0;
        if (t < ntask) {
            if (EDGE) {
                s = tsz[t];
            } else {
                s = rowptr[t + 1] - rowptr[t];
            }
            if (s >= need) c = class_of(s);
        }
        for (int q = 0; q < NCLASS; ++q) {
            uint64_t bal = __ballot(c == q);
            if (!bal) continue;
            int leader = __ffsll((unsigned long long)bal) - 1;
            if (!scatter) {
                uint32_t mx = (c == q) ? s : 0;
                for (int d = 32; d >= 1; d >>= 1) mx = max(mx, (uint32_t)__shfl_xor(mx, d, 64));
                if (lane == leader) {
                    atomicAdd(&cls_cnt[q], (uint32_t)__popcll(bal));
                    atomicMax(&cls_max[q], mx);
                }
            } else {
                uint32_t b0 = 0;
                if (lane == leader) b0 = atomicAdd(&cls_pos[q], (uint32_t)__popcll(bal));
                b0 = __shfl(b0, leader, 64);
                if (c == q) out[cls_off[q] + b0 + (uint32_t)__popcll(bal & lanemask_lt())] = (uint32_t)t;
            }
        }
    }
}

// ---- host ----------------------------------------------------------------------
template <bool PIVOT, bool EDGE, bool GLOBAL>
static const void *kernel_ptr()
{
    return reinterpret_cast<const void *>(&k_count<PIVOT, EDGE, GLOBAL>);
}

static const void *pick_kernel(bool pivot, bool edge, bool global)
{
    if (pivot) {
        if (edge) return global ? kernel_ptr<true, true, true>() : kernel_ptr<true, true, false>();
        return global ? kernel_ptr<true, false, true>() : kernel_ptr<true, false, false>();
    }
    if (edge) return global ? kernel_ptr<false, true, true>() : kernel_ptr<false, true, false>();
    return global ? kernel_ptr<false, false, true>() : kernel_ptr<false, false, false>();
}

// exact C(n, r) for n <= nmax, r <= rmax; ovf[i] = 1 when it exceeds 2^64-1
static void binom_table(uint32_t nmax, uint32_t rmax, std::vector<uint64_t> &val, std::vector<uint8_t> &ovf)
{
    const uint32_t st = rmax + 1;
    val.assign((size_t)(nmax + 1) * st, 0);
    ovf.assign((size_t)(nmax + 1) * st, 0);
    for (uint32_t n = 0; n <= nmax; ++n) {
        val[(size_t)n * st] = 1;
        for (uint32_t r = 1; r <= rmax && r <= n; ++r) {
            size_t i = (size_t)n * st + r, a = (size_t)(n - 1) * st + r - 1, b = (size_t)(n - 1) * st + r;
            uint64_t s = val[a] + val[b];
            ovf[i] = ovf[a] | ovf[b] | (s < val[a]);
            val[i] = s;
        }
    }
}

struct ClassPlan {
    uint32_t count = 0, maxsize = 0, offset = 0;
};

int count_impl(const kc_oriented *o, const kc_query *q, kc_result *r)
{
    DeviceCtx *ctx;
    int rc = device_ctx(&ctx);
    if (rc) return rc;
    kc_oriented *om = const_cast<kc_oriented *>(o);
    hipStream_t s = om->stream;
    const uint32_t k = q->k;
    const bool pivot = q->algo == KC_ALGO_PIVOT, edge = q->gran == KC_GRAN_EDGE;
    const uint32_t P = q->part_size;
    const uint32_t ntask = edge ? (uint32_t)o->m : o->n;
    const uint32_t need = edge ? k - 2 : k - 1;  // k >= 3 here

    // --- task classes (two passes: histogram + max, then scatter) ---
    // meta: cnt[6] max[6] pos[6] off[6] qhead[6] ovf[1] pad, acc[6*4] (u64)
    const size_t meta_u32 = 64;
    const size_t meta_bytes = meta_u32 * 4 + sizeof(uint64_t) * NCLASS * ACC;
    if ((rc = om->s_meta.ensure(meta_bytes))) return rc;
    uint32_t *m32 = om->s_meta.as<uint32_t>();
    uint32_t *cls_cnt = m32, *cls_max = m32 + 8, *cls_pos = m32 + 16, *cls_off = m32 + 24, *qhead = m32 + 32,
             *dovf = m32 + 40;
    uint64_t *dacc = reinterpret_cast<uint64_t *>(m32 + meta_u32);
    KC_HIP(hipMemsetAsync(om->s_meta.p, 0, meta_bytes, s));
    int cg = (int)std::min<uint64_t>(((uint64_t)ntask + 255) / 256, (uint64_t)ctx->cu * 8);
    if (cg < 1) cg = 1;
    uint32_t *tsz = nullptr;
    if (edge && ntask > 0) {
        if ((rc = om->s_tsz.ensure(sizeof(uint32_t) * (size_t)ntask))) return rc;
        tsz = om->s_tsz.as<uint32_t>();
        int tg = (int)std::min<uint64_t>(((uint64_t)ntask + 255) / 256, (uint64_t)ctx->cu * 16);
        hipLaunchKernelGGL(k_tsize, dim3(tg), dim3(256), 0, s, o->rowptr.as<uint32_t>(), o->col.as<uint32_t>(),
                           o->src.as<uint32_t>(), ntask, tsz);
    }
    auto classify = [&](int scatter) {
        if (edge)
            hipLaunchKernelGGL(k_classify<true>, dim3(cg), dim3(256), 0, s, o->rowptr.as<uint32_t>(), tsz, ntask,
                               need, cls_cnt, cls_max, cls_pos, om->s_tasks.as<uint32_t>(), cls_off, scatter);
        else
            hipLaunchKernelGGL(k_classify<false>, dim3(cg), dim3(256), 0, s, o->rowptr.as<uint32_t>(), tsz, ntask,
                               need, cls_cnt, cls_max, cls_pos, om->s_tasks.as<uint32_t>(), cls_off, scatter);
    };
    if (ntask > 0) classify(0);
    KC_HIP(hipGetLastError());
    uint32_t h[16];
    KC_HIP(hipMemcpyAsync(h, m32, sizeof h, hipMemcpyDeviceToHost, s));
    KC_HIP(hipStreamSynchronize(s));
    ClassPlan plan[NCLASS];
    uint32_t tot = 0, maxsize = 0;
    for (int c = 0; c < NCLASS; ++c) {
        plan[c].count = h[c];
        plan[c].maxsize = h[8 + c];
        plan[c].offset = tot;
        tot += h[c];
        maxsize = std::max(maxsize, h[8 + c]);
    }
    if ((rc = om->s_tasks.ensure(sizeof(uint32_t) * ((size_t)tot + 1)))) return rc;
    uint32_t hoff[NCLASS];
    for (int c = 0; c < NCLASS; ++c) hoff[c] = plan[c].offset;
    KC_HIP(hipMemcpyAsync(cls_off, hoff, sizeof hoff, hipMemcpyHostToDevice, s));
    if (tot > 0) classify(1);
    KC_HIP(hipGetLastError());

    // --- binomial table (pivoting) ---
    const uint32_t bstride = k + 1;
    if (pivot) {
        std::vector<uint64_t> bv;
        std::vector<uint8_t> bo;
        binom_table(maxsize + 1, k, bv, bo);
        size_t nb = bv.size();
        if ((rc = om->s_binom.ensure(nb * 9 + 16))) return rc;
        KC_HIP(hipMemcpyAsync(om->s_binom.p, bv.data(), nb * 8, hipMemcpyHostToDevice, s));
        KC_HIP(hipMemcpyAsync(om->s_binom.as<uint8_t>() + nb * 8, bo.data(), nb, hipMemcpyHostToDevice, s));
        KC_HIP(hipStreamSynchronize(s));  // host vectors go out of scope
    }
    const size_t nbin = (size_t)(maxsize + 2) * bstride;

    // --- launches, largest class first ---
    hipEvent_t ev[NCLASS + 1];
    for (auto &e : ev) hipEventCreate(&e);
    float cls_ms[NCLASS] = {0};
    bool launched[NCLASS] = {false};
    uint32_t variant[NCLASS] = {0};
    hipEventRecord(ev[NCLASS], s);
    const size_t lds_cap = ctx->lds_per_block;
    for (int c = NCLASS - 1; c >= 0; --c) {
        if (plan[c].count == 0) continue;
        CountArgs a{};
        a.rowptr = o->rowptr.as<uint32_t>();
        a.col = o->col.as<uint32_t>();
        a.src = o->src.as<uint32_t>();
        a.tasks = om->s_tasks.as<uint32_t>() + plan[c].offset;
        a.ntasks = plan[c].count;
        a.k = k;
        a.psize = P;
        a.binom = pivot ? om->s_binom.as<uint64_t>() : nullptr;
        a.binom_ovf = pivot ? om->s_binom.as<uint8_t>() + nbin * 8 : nullptr;
        a.bstride = bstride;
        a.qhead = qhead + c;
        a.acc = dacc + ACC * c;
        a.ovf = dovf;
        bool global = (c == NCLASS - 1);
        uint32_t rows_cap = global ? plan[c].maxsize : CLASS_ROWS[c];
        a.rmax = rows_cap;
        a.wmax = (rows_cap + 63) / 64;
        a.levels = pivot ? k + 3 : std::max<uint32_t>(1, k - 2);
        uint32_t np = 64 / P;
        Layout L = make_layout(a.rmax, a.wmax, a.levels, np, pivot);
        if (!global && !pivot) {
            // fewer partitions when the stacks would not fit in LDS
            while (np > 1 && L.total * 8 > lds_cap) {
                np >>= 1;
                L = make_layout(a.rmax, a.wmax, a.levels, np, pivot);
            }
        }
        if (!global && L.total * 8 > lds_cap) {
            global = true;  // does not fit: same layout in HBM
        }
        a.nparts = np;
        const void *kern = pick_kernel(pivot, edge, global);
        if (pivot && !global && c <= 2) {
            // register-bitmap SCT: rows live in VGPRs, LDS only for the build
            static const void *const reg[2][3] = {
                {reinterpret_cast<const void *>(&k_pivot_reg<false, 1>), reinterpret_cast<const void *>(&k_pivot_reg<false, 2>),
                 reinterpret_cast<const void *>(&k_pivot_reg<false, 4>)},
                {reinterpret_cast<const void *>(&k_pivot_reg<true, 1>), reinterpret_cast<const void *>(&k_pivot_reg<true, 2>),
                 reinterpret_cast<const void *>(&k_pivot_reg<true, 4>)}};
            kern = reg[edge ? 1 : 0][c];
            L = make_layout(a.rmax, a.wmax, 0, 0, true);
        }
        size_t lds = global ? 0 : L.total * 8;
        if (!global) KC_HIP(hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        int per_cu = 0;
        KC_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 64, lds));
        if (per_cu < 1) per_cu = 1;
        uint64_t blocks = (uint64_t)per_cu * ctx->cu;
        a.chunk = (c == 0 && plan[c].count > blocks * 64) ? 4 : 1;
        blocks = std::min<uint64_t>(blocks, ((uint64_t)plan[c].count + a.chunk - 1) / a.chunk);
        if (global) {
            a.gscr_words = L.total;
            uint64_t cap = (2ull << 30) / (L.total * 8 + 1);
            blocks = std::max<uint64_t>(1, std::min<uint64_t>(blocks, cap));
            if ((rc = om->s_gscr.ensure(blocks * L.total * 8))) {
                for (auto &e : ev) hipEventDestroy(e);
                return rc;
            }
            a.gscr = om->s_gscr.as<uint64_t>();
        }
        void *args[] = {&a};
        KC_HIP(hipLaunchKernel(kern, dim3((unsigned)blocks), dim3(64), args, lds, s));
        variant[c] = global ? 3u : (pivot && c <= 2) ? 2u : 1u;
        hipEventRecord(ev[c], s);
        launched[c] = true;
    }
    uint64_t hacc[NCLASS * ACC];
    uint32_t hovf = 0;
    KC_HIP(hipMemcpyAsync(hacc, dacc, sizeof hacc, hipMemcpyDeviceToHost, s));
    KC_HIP(hipMemcpyAsync(&hovf, dovf, sizeof hovf, hipMemcpyDeviceToHost, s));
    hipError_t se = hipStreamSynchronize(s);
    if (se != hipSuccess) {
        for (auto &e : ev) hipEventDestroy(e);
        return fail(KC_ERR_DEVICE, "counting kernels failed: %s", hipGetErrorString(se));
    }
    // per-class device times: each class starts where the previous launch ended
    hipEvent_t prev = ev[NCLASS];
    float total_ms = 0;
    for (int c = NCLASS - 1; c >= 0; --c) {
        if (!launched[c]) continue;
        hipEventElapsedTime(&cls_ms[c], prev, ev[c]);
        total_ms += cls_ms[c];
        prev = ev[c];
    }
    for (auto &e : ev) hipEventDestroy(e);

    std::memset(r, 0, sizeof *r);
    int wall_khz = 100000;
    hipDeviceGetAttribute(&wall_khz, hipDeviceAttributeWallClockRate, ctx->dev);
    if (wall_khz <= 0) wall_khz = 100000;
    uint64_t cnt = 0;
    bool ov = hovf != 0;
    int dom = -1;
    for (int c = 0; c < NCLASS; ++c) {
        uint64_t v = hacc[ACC * c];
        if (cnt + v < cnt) ov = true;
        cnt += v;
        r->alg_bytes += hacc[ACC * c + 1];
        r->tasks += hacc[ACC * c + 2];
        r->tree_nodes += hacc[ACC * c + 3];
        if (launched[c] && (dom < 0 || cls_ms[c] > cls_ms[dom])) dom = c;
        r->class_ms[c] = cls_ms[c];
        r->class_bytes[c] = hacc[ACC * c + 1];
        r->class_tasks[c] = hacc[ACC * c + 2];
        r->class_kernel[c] = variant[c];
        r->class_max_task_ms[c] = (double)hacc[ACC * c + 4] / (double)wall_khz;
        r->class_task_ms_sum[c] = (double)hacc[ACC * c + 5] / (double)wall_khz;
    }
    r->count = cnt;
    r->overflow = ov ? 1 : 0;
    r->algo = q->algo;
    r->gran = q->gran;
    r->part_size = P;
    r->count_ms = total_ms;
    if (dom >= 0) {
        r->dominant_ms = cls_ms[dom];
        r->dominant_bytes = hacc[ACC * dom + 1];
        r->dominant_class = (uint32_t)dom;
    }
    return ov ? fail(KC_ERR_OVERFLOW, "k-clique count exceeds 2^64-1") : KC_OK;
}

}  // namespace kc

using namespace kc;

extern "C" int kc_count(const kc_oriented *o, const kc_query *q, kc_result *r)
{
    if (!o || !q || !r) return fail(KC_ERR_ARG, "kc_count: null argument");
    if (q->k < 1 || q->k > KC_MAX_K) return fail(KC_ERR_ARG, "kc_count: k=%u outside 1..%d", q->k, KC_MAX_K);
    if (q->algo > 1 || q->gran > 1) return fail(KC_ERR_ARG, "kc_count: bad algo/gran");
    uint32_t P = q->part_size;
    if (P == 0 || P > 64 || (P & (P - 1))) return fail(KC_ERR_ARG, "kc_count: partition size %u", P);
    hipSetDevice(o->dev);
    kc_query qq = *q;
    if (q->auto_tune) {
        // Measured on gfx950: one lane per candidate row / per DFS partition
        // (P = 1) is fastest for every size class, and pivoting with
        // closed-form tails is never slower than the orientation DFS;
        // edge tasks balance the heavy-tailed degree distributions.
        qq.algo = KC_ALGO_PIVOT;
        qq.gran = o->m > 0 ? KC_GRAN_EDGE : KC_GRAN_NODE;
        qq.part_size = 1;
    }
    q = &qq;
    P = q->part_size;
    if (q->k <= 2) {
        // closed forms: every vertex is a 1-clique, every edge a 2-clique
        std::memset(r, 0, sizeof *r);
        r->count = q->k == 1 ? o->n : o->m;
        r->algo = q->algo;
        r->gran = q->gran;
        r->part_size = P;
        r->tasks = q->gran == KC_GRAN_EDGE ? o->m : o->n;
        return KC_OK;
    }
    return count_impl(o, q, r);
}
