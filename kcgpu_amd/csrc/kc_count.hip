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
//      pivoting: one SCT tree per wave (per block of waves above 256
//        vertices); the pivot is the smallest index among the candidates j
//        maximising |X & row_j|; leaves add C(pivots, k - held) from an exact
//        uint64 table with overflow bits.  Branches on non-pivot vertices are
//        taken first and the pivot branch last as a tail call, so the stack
//        never holds more than k - held + 2 levels (the reference keeps
//        maxDeg levels in HBM).  Subtrees are re-encoded into one-, two- or
//        four-word rows held in registers once their candidate sets shrink,
//        and the last levels close in form (edges; triangles by lane walks;
//        3- and 4-cliques of near-cliques from the complement).
// Sizes: tasks are binned by their exact list size (64, 128, 256 vertices in
// registers, 512, 1024 in LDS, larger in an HBM slice), each class cut into
// slices drained largest first.
//
// Also here: the serving loop's resident launches (k_pivot_open and its host
// side, OpenEngine), which take tasks from queues the host extends while they
// run.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include "kc_device.h"

namespace kc {

constexpr int NCLASS = 6;
constexpr uint32_t CLASS_ROWS[NCLASS] = {64, 128, 256, 512, 1024, 0xFFFFFFFFu};
constexpr uint32_t INF32 = 0xFFFFFFFFu;
constexpr int ACC = 8;  // per-class accumulator slots
constexpr uint32_t PIVOT_TOP2 = 256;  // largest pivoting task of the four-word register kernel

// one query of a launch: the oriented graph it runs on and its k
struct QDesc {
    const uint32_t *rowptr, *col, *src;
    uint32_t k, gran, sample, offset;  // tasks with id % sample == offset run
    uint32_t psize;                    // lanes per partition of this query
};
static_assert(sizeof(QDesc) % 8 == 0, "QDesc is read as 64-bit words");

// the graph and k of the task a wave is working on
struct Cur {
    const uint32_t *rowptr, *col, *src;
    uint32_t k, gran, psize;
};

struct CountArgs {
    const QDesc *qd;
    const uint64_t *tasks;  // (query << 32) | vertex-or-edge id
    uint32_t ntasks;
    uint32_t nparts;        // partitions per wave the orientation DFS stacks are laid out for
    uint32_t chunk;         // tasks per queue grab
    uint32_t rmax;          // row capacity of the workspace
    uint32_t wmax;          // words per row capacity
    uint32_t levels;        // stack levels
    const uint64_t *binom;
    const uint8_t *binom_ovf;
    uint32_t bstride;       // binom[n * bstride + r]
    uint32_t *qhead;
    uint64_t *acc;          // [0] count [1] bytes [2] tasks [3] tree nodes [4] max task ticks [5] sum task ticks
    uint64_t *qcount;       // per query
    uint32_t *qovf;         // per query
    uint64_t *gscr;
    uint64_t gscr_words;
    uint64_t *local;        // per-vertex counts of the (single) query, or null
    uint32_t chain;         // wide kernel: pivot-chain levels walked before the entries
};

// ---- workspace layout (u64 units) -----------------------------------------
// Per-vertex counting (LOCAL) adds, for pivoting, the held and pivot sets of
// the current node and of every stack level, and for both traversals a u64
// accumulator per task vertex (lacc) and the DFS path of each partition.
struct Layout {
    uint64_t misc, list, offs, starts, rows, stk, x, scal, lacc, path, total;
};

__host__ __device__ inline Layout make_layout(uint32_t rmax, uint32_t wmax, uint32_t levels, uint32_t nparts,
                                              bool pivot, bool local = false)
{
    Layout L;
    L.misc = 0;
    L.list = L.misc + 8;
    L.starts = L.list + (rmax + 1) / 2;  // list and starts are dead once the rows are built
    L.offs = L.starts + (rmax + 1) / 2;
    L.rows = L.offs + (rmax + 2) / 2;
    L.stk = L.rows + (uint64_t)rmax * wmax;
    const uint64_t sets = local ? 4 : 2;  // per level: candidates, branches (+ held, pivots)
    if (pivot) {
        L.x = L.stk + (uint64_t)levels * sets * wmax;
        L.scal = L.x + (local ? 3 : 1) * (uint64_t)wmax;  // X (+ current held, pivots)
        L.lacc = L.scal + (uint64_t)levels * 2;
        L.path = L.lacc + (local ? rmax : 0);
        L.total = L.path;
    } else {
        L.x = L.scal = L.stk + (uint64_t)nparts * levels * 2 * wmax;
        L.lacc = L.x;
        L.path = L.lacc + (local ? rmax : 0);
        L.total = L.path + (local ? ((uint64_t)nparts * levels + 1) / 2 : 0);
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
__device__ __attribute__((always_inline)) void build_rows(const Cur &g, uint32_t *list, uint32_t *offs,
                                                         uint32_t *starts, uint64_t *rows, uint32_t R, uint32_t W,
                                                         uint64_t &bytes)
{
    const int lane = lane_id();
    for (uint32_t i = lane; i < R * W; i += 64) rows[i] = 0;
    uint32_t carry = 0;
    for (uint32_t base = 0; base < R; base += 64) {
        uint32_t i = base + lane, d = 0;
        if (i < R) {
            uint32_t w = list[i];
            uint32_t s = g.rowptr[w];
            d = g.rowptr[w + 1] - s;
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
    uint32_t lo = 0;  // row of entry f: the last i with offs[i] <= f (f only grows)
    for (uint32_t f = lane; f < total; f += 64) {
        while (offs[lo + 1] <= f) ++lo;
        uint32_t x = g.col[starts[lo] + (f - offs[lo])];
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
// LOCAL also adds, per found clique, 1 to each member in lacc (the task
// vertices; the root is credited with the task's count by the caller):
// partitions share lacc, so those adds are LDS atomics.
template <bool LOCAL>
__device__ uint64_t orient_count(const CountArgs &a, uint32_t P, const uint64_t *rows, uint64_t *stk,
                                 uint32_t *next_row, uint32_t R, uint32_t W, uint32_t r, uint64_t &nodes, uint32_t &ovf,
                                 uint64_t *lacc, uint32_t *path)
{
    const int lane = lane_id();
    uint64_t cnt = 0;
    if (r == 2) {
        for (uint32_t i = lane; i < R * W; i += 64) {
            const uint64_t x = rows[i];
            cnt += (uint64_t)__popcll(x);
            if (LOCAL && x) {
                const uint32_t row = i / W, wbase = (i % W) * 64;
                atomicAdd((unsigned long long *)&lacc[row], (unsigned long long)__popcll(x));
                for (uint64_t b = x; b; b &= b - 1)
                    atomicAdd((unsigned long long *)&lacc[wbase + (uint32_t)__builtin_ctzll(b)], 1ull);
            }
        }
        return cnt;
    }
    // P lanes per partition (the task's query); the stacks hold a.nparts partitions
    const uint32_t NP = 64 / P < a.nparts ? 64 / P : a.nparts;
    const uint32_t part = lane / P, lip = lane % P;
    if (part >= NP) return 0;
    const uint32_t wmax = a.wmax;
    uint64_t *base = stk + (uint64_t)part * a.levels * 2 * wmax;
    uint32_t *mypath = LOCAL ? path + (uint64_t)part * a.levels : nullptr;
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
            if (LOCAL && lip == 0) mypath[t] = b;
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
                if (LOCAL && c) {
                    // clique = root + j + path[0..t] + one vertex of Cn
                    for (uint32_t w = lip; w < W; w += P)
                        for (uint64_t x = Cn[w]; x; x &= x - 1)
                            atomicAdd((unsigned long long *)&lacc[w * 64 + (uint32_t)__builtin_ctzll(x)], 1ull);
                    if (lip == 0) {
                        atomicAdd((unsigned long long *)&lacc[j], (unsigned long long)c);
                        for (int l = 0; l <= t; ++l)
                            atomicAdd((unsigned long long *)&lacc[mypath[l]], (unsigned long long)c);
                    }
                }
            } else if (c >= remc) {
                ++t;
            }
        }
    }
    if (lip == 0) nodes += my_nodes;
    return cnt;
}

// per-lane gather: lane l gets v of lane src(l)
__device__ inline uint64_t bperm64(uint64_t v, uint32_t src)
{
    const uint32_t lo = (uint32_t)__builtin_amdgcn_ds_bpermute((int)(src << 2), (int)(uint32_t)v);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_ds_bpermute((int)(src << 2), (int)(uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}

// A value every lane holds alike, moved to scalar registers: the compiler then
// treats what is derived from it as uniform (scalar ALU, scalar branches)
// instead of masking the wave lane by lane.
__device__ inline uint32_t uni32(uint32_t v) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)v); }

__device__ inline uint64_t uni64(uint64_t v)
{
    return ((uint64_t)uni32((uint32_t)(v >> 32)) << 32) | uni32((uint32_t)v);
}

__device__ inline uint64_t rdlane64(uint64_t v, uint32_t src)
{
    uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, (int)src);
    uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), (int)src);
    return ((uint64_t)hi << 32) | lo;
}

// ---- closed forms over a one-word candidate set X (row = the lane's row) ----
// All of them run with every lane active: rows of other lanes come in with
// ds_bpermute, and loops run while any lane still has work (ballot).

// triangles of G[X] through the complement (Goodman):
//   t(G) = C(|X|,3) - 1/2 sum_x d_x (|X|-1-d_x) - t(complement),
// each lane walking its complement neighbours above it (cmpl, at most cmax)
__device__ inline uint32_t tri_complement(uint64_t row, uint64_t cmpl, uint32_t cmax, bool mem, uint32_t v, uint32_t cx)
{
    uint32_t gd = mem ? v * (cx - 1 - v) : 0u, tb = 0;
    uint64_t rest = cmpl;
    for (uint32_t it = 0; it < cmax; ++it) {
        const uint32_t y = rest ? (uint32_t)__builtin_ctzll(rest) : 0u;
        const uint64_t ry = bperm64(row, y);
        if (rest) {
            const uint64_t ay = y == 63 ? 0ull : (~0ull << (y + 1));
            tb += (uint32_t)__popcll(cmpl & ~ry & ay);
            rest &= rest - 1;
        }
    }
    gd = (uint32_t)__builtin_amdgcn_readfirstlane((int)__ockl_wfred_add_u32(gd));
    tb = (uint32_t)__builtin_amdgcn_readfirstlane((int)__ockl_wfred_add_u32(tb));
    return cx * (cx - 1) * (cx - 2) / 6 - gd / 2 - tb;
}

// triangles of G[X], each i < j < l counted at i over a uniform walk of X
__device__ inline uint32_t tri_direct(uint64_t X, uint64_t row, uint64_t lane_above)
{
    const uint32_t lane = (uint32_t)lane_id();
    uint32_t tri = 0;
    const uint64_t mine = X & row & lane_above;
    for (uint64_t bits = X; bits; bits &= bits - 1) {
        const uint32_t il = (uint32_t)__builtin_ctzll(bits);
        const uint64_t above = il == 63 ? 0ull : (~0ull << (il + 1));
        const uint64_t Y = X & rdlane64(row, il) & above;
        if ((Y >> lane) & 1ull) tri += (uint32_t)__popcll(Y & mine);
    }
    return (uint32_t)__builtin_amdgcn_readfirstlane((int)__ockl_wfred_add_u32(tri));
}


// |E|, triangles and 4-cliques of G[X] from its complement H (the non-edges
// inside X), by inclusion-exclusion over the edge sets F of H that span at
// most four vertices, sum_F (-1)^|F| C(n - |V(F)|, 4 - |V(F)|):
//   c4 = C(n,4) - m C(n-2,2) + C(m,2) + (n-4) P3 - (n-3) K3 - P4 - S3 + C4 + paw - dia + K4
// where, all as subgraphs of H: m edges, P3 = sum_x C(d_x,2) two-edge paths
// (C(m,2) - P3 is the number of two-edge matchings), K3 triangles,
// P4 = sum_{xy in E} (d_x-1)(d_y-1) - 3 K3 three-edge paths, S3 = sum_x C(d_x,3)
// three-edge stars, C4 = 1/2 sum_{x<y} C(c_xy,2) four-cycles (c_xy = common
// neighbours), paw = sum_x t_x (d_x-2) triangles with a pendant edge, dia =
// sum_{xy in E} C(c_xy,2) four-cycles with a chord, K4 four-cliques; and
// Goodman's c3 = C(n,3) - 1/2 sum_x d_x (n-1-d_x) - K3.
// Lane x holds its complement row Hx.  One uniform walk over u in X (row u by
// v_readlane) hands c_ux = |Hu & Hx| to every lane, which keeps
//   q  = sum_u c_ux^2,   qE = sum_{u in Hx} c_ux^2,   t2 = sum_{u in Hx} c_ux = 2 t_x,
//   sd = sum_{u in Hx} d_u  (= sum_u c_ux),
// so that, summed over lanes, 8 C4 = sum (q - d^2) - (sd - d), 2 P4 + 6 K3 =
// sum (d-1)(sd-d), 4 dia = sum qE - t2.  The lane shares are summed scaled by
// 24 so each is an integer.  A K4 of H has every vertex in at least three
// H-triangles, so the K4 walk (over the H-edges u < v, counting the H-edges
// inside Hu & Hv above v) only visits such vertices and is empty in most
// near-cliques.
struct CliqueCounts {
    uint32_t e, c3, c4;
};

__device__ inline CliqueCounts k34_complement(uint64_t X, uint64_t row, bool mem, uint32_t cx)
{
    const uint32_t lane = (uint32_t)lane_id();
    const uint64_t Hl = mem ? (X & ~row & ~(1ull << lane)) : 0ull;
    const int32_t d = (int32_t)__popcll(Hl), n = (int32_t)cx;
    uint32_t q = 0, qE = 0, t2 = 0, sd = 0;
    if ((X >> 32) == 0) {
        // X within vertices 0..31: lanes l and l+32 both hold vertex l (32-bit
        // rows) and walk two members u per step, one per half of the wave
        const uint32_t lo = lane & 31, X32 = (uint32_t)X;
        const bool upper = lane >= 32;
        const uint32_t row32 = (uint32_t)__shfl((int)(uint32_t)row, (int)lo, 64);
        const uint32_t H32 = ((X32 >> lo) & 1u) ? (X32 & ~row32 & ~(1u << lo)) : 0u;
        for (uint32_t bits = X32; bits;) {
            const uint32_t ua = (uint32_t)__builtin_ctz(bits);
            bits &= bits - 1;
            uint32_t Hb = 0;
            if (bits) {
                const uint32_t ub = (uint32_t)__builtin_ctz(bits);
                bits &= bits - 1;
                Hb = X32 & ~(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)row, (int)ub) & ~(1u << ub);
            }
            const uint32_t Ha = X32 & ~(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)row, (int)ua) & ~(1u << ua);
            const uint32_t Hu = upper ? Hb : Ha;
            const uint32_t du = upper ? (uint32_t)__popc(Hb) : (uint32_t)__popc(Ha);
            const uint32_t c = (uint32_t)__popc(H32 & Hu), c2 = c * c;
            const uint32_t in = (Hu >> lo) & 1u;  // the lane's vertex is a complement neighbour of u
            q += c2;
            qE += in * c2;
            t2 += in * c;
            sd += in * du;
        }
        q += (uint32_t)__shfl_xor((int)q, 32, 64);
        qE += (uint32_t)__shfl_xor((int)qE, 32, 64);
        t2 += (uint32_t)__shfl_xor((int)t2, 32, 64);
        sd += (uint32_t)__shfl_xor((int)sd, 32, 64);
        if (upper) q = qE = t2 = sd = 0;  // (Hl, d are 0 there: X has no member above 31)
    } else {
        for (uint64_t bits = X; bits; bits &= bits - 1) {
            const uint32_t u = (uint32_t)__builtin_ctzll(bits);
            const uint64_t Hu = X & ~rdlane64(row, u) & ~(1ull << u);
            const uint32_t du = (uint32_t)__popcll(Hu);
            const uint32_t c = (uint32_t)__popcll(Hl & Hu), c2 = c * c;
            const uint32_t in = (uint32_t)(Hu >> lane) & 1u;
            q += c2;
            qE += in * c2;
            t2 += in * c;
            sd += in * du;
        }
    }
    const int32_t t = (int32_t)(t2 / 2);  // H-triangles through the lane's vertex
    // K4 of H among the vertices in >= 3 H-triangles
    int32_t k4acc = 0;
    const uint64_t T3 = __ballot(t >= 3);
    if (__popcll(T3) >= 4) {
        const uint64_t lane_above = lane == 63 ? 0ull : (~0ull << (lane + 1));
        for (uint64_t ub = T3; ub; ub &= ub - 1) {
            const uint32_t u = (uint32_t)__builtin_ctzll(ub);
            const uint64_t Hu = T3 & ~rdlane64(row, u) & (u == 63 ? 0ull : (~0ull << (u + 1)));
            for (uint64_t vb = Hu; vb; vb &= vb - 1) {
                const uint32_t v = (uint32_t)__builtin_ctzll(vb);
                const uint64_t S = Hu & ~rdlane64(row, v) & (v == 63 ? 0ull : (~0ull << (v + 1)));
                if (__popcll(S) >= 2 && ((S >> lane) & 1ull)) k4acc += (int32_t)__popcll(S & Hl & lane_above);
            }
        }
    }
    const int32_t iq = (int32_t)q, iqE = (int32_t)qE, isd = (int32_t)sd, it2 = (int32_t)t2;
    const int32_t c4l = 12 * (n - 4) * d * (d - 1) - 8 * (n - 3) * t - 12 * (d - 1) * (isd - d) + 24 * t -
                        4 * d * (d - 1) * (d - 2) + 3 * ((iq - d * d) - (isd - d)) + 24 * t * (d - 2) -
                        6 * (iqE - it2) + 24 * k4acc;
    const int32_t c3l = -3 * d * (n - 1 - d) - 2 * t;
    const int32_t s_d = (int32_t)__builtin_amdgcn_readfirstlane((int)__ockl_wfred_add_u32((uint32_t)d));
    const int32_t s_4 = (int32_t)__builtin_amdgcn_readfirstlane((int)__ockl_wfred_add_u32((uint32_t)c4l));
    const int32_t s_3 = (int32_t)__builtin_amdgcn_readfirstlane((int)__ockl_wfred_add_u32((uint32_t)c3l));
    const int32_t m = s_d / 2;
    const int32_t c4_24 = n * (n - 1) * (n - 2) * (n - 3) - 12 * m * (n - 2) * (n - 3) + 12 * m * (m - 1) + s_4;
    const int32_t c3_6 = n * (n - 1) * (n - 2) + s_3;
    CliqueCounts r;
    r.e = (uint32_t)(n * (n - 1) / 2 - m);
    r.c3 = (uint32_t)(c3_6 / 6);
    r.c4 = (uint32_t)(c4_24 / 24);
    return r;
}

__device__ inline uint64_t mul_chk(uint64_t x, uint64_t y, uint32_t &ovf)
{
    uint64_t r;
    if (__builtin_mul_overflow(x, y, &r)) ovf = 1;
    return r;
}

__device__ inline uint64_t binom_chk(const CountArgs &a, uint32_t n, uint32_t r, uint32_t &ovf)
{
    if (r > n) return 0;
    const uint64_t idx = (uint64_t)n * a.bstride + r;
    if (a.binom_ovf[idx]) ovf = 1;
    return a.binom[idx];
}

// k-cliques of a subtree at h = k-3: C(p,3) + C(p,2)|X| + p|E(X)| + t(X)
__device__ inline uint64_t close_k3(const CountArgs &a, uint32_t p, uint32_t cx, uint64_t e, uint64_t tri,
                                    uint32_t &ovf)
{
    uint64_t r = 0;
    add_checked(r, binom_chk(a, p, 3, ovf), ovf);
    add_checked(r, mul_chk(binom_chk(a, p, 2, ovf), cx, ovf), ovf);
    add_checked(r, mul_chk(p, e, ovf), ovf);
    add_checked(r, tri, ovf);
    return r;
}

// SCT of a subproblem of at most 64 vertices held one row per lane: lane i
// keeps row i (bit j <=> vertex j of the subproblem is adjacent), X0 is the
// candidate set, h0/p0 the held and pivot counts above it.  Every set is one
// uniform 64-bit word, a score is one AND + popcount, a branch row is one
// v_readlane pair, and the stack lives across the lanes (lane t = level t).
// Nodes with h >= k-3 close in form as in pivot_reg, nodes with h = k-4 by
// k34_complement.
__device__ uint64_t sct64(const CountArgs &a, uint32_t k, uint64_t row, uint64_t X0, uint32_t h0, uint32_t p0,
                          uint64_t &nodes, uint32_t &ovf)
{
    const uint32_t lane = (uint32_t)lane_id();
    const uint64_t lane_above = lane == 63 ? 0ull : (~0ull << (lane + 1));
    k = uni32(k);
    uint64_t X = uni64(X0), stkC = 0, stkB = 0;
    uint32_t stkM = 0;
    int t = -1;
    bool pending = true;
    uint32_t h = uni32(h0), p = uni32(p0);
    uint64_t acc = 0, my_nodes = 0;
    for (;;) {
        if (pending) {
            pending = false;
            ++my_nodes;
            const uint32_t cx = (uint32_t)__popcll(X);
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
                const bool mem = (X >> lane) & 1ull;
                const uint32_t v = mem ? (uint32_t)__popcll(X & row) : 0u;
                if (h + 3 >= k) {
                    const uint32_t e2 = (uint32_t)__builtin_amdgcn_readfirstlane((int)__ockl_wfred_add_u32(v));
                    const uint64_t pp = p;
                    if (h + 2 == k) {
                        add_checked(acc, pp * (pp - (pp ? 1 : 0)) / 2, ovf);
                        add_checked(acc, pp * cx, ovf);
                        add_checked(acc, e2 / 2, ovf);
                    } else {
                        // triangles of G[X]: through the complement when X is nearly a
                        // clique, else every triangle i < j < l counted at i
                        const uint64_t cmpl = mem ? (X & ~row & lane_above) : 0ull;
                        const uint32_t cmax =
                            (uint32_t)__builtin_amdgcn_readfirstlane((int)__ockl_wfred_max_u32((uint32_t)__popcll(cmpl)));
                        const uint32_t tri = (3 * cmax < 2 * cx) ? tri_complement(row, cmpl, cmax, mem, v, cx)
                                                                 : tri_direct(X, row, lane_above);
                        add_checked(acc, pp * (pp - (pp ? 1 : 0)) * (pp >= 2 ? pp - 2 : 0) / 6, ovf);
                        add_checked(acc, pp * (pp - (pp ? 1 : 0)) / 2 * cx, ovf);
                        add_checked(acc, pp * (e2 / 2), ovf);
                        add_checked(acc, tri, ovf);
                    }
                } else if (h + 4 == k) {
                    // h = k-4: the subtree holds
                    // C(p,4) + C(p,3)|X| + C(p,2)|E(X)| + p c3(X) + c4(X) k-cliques
                    const CliqueCounts cc = k34_complement(X, row, mem, cx);
                    add_checked(acc, binom_chk(a, p, 4, ovf), ovf);
                    add_checked(acc, mul_chk(binom_chk(a, p, 3, ovf), cx, ovf), ovf);
                    add_checked(acc, mul_chk(binom_chk(a, p, 2, ovf), cc.e, ovf), ovf);
                    add_checked(acc, mul_chk(p, cc.c3, ovf), ovf);
                    add_checked(acc, cc.c4, ovf);
                } else {
                    // (score + 1, 63 - index): the max names the smallest index of the largest score
                    uint32_t key = mem ? (((v + 1) << 8) | (63u - lane)) : 0u;
                    key = (uint32_t)__builtin_amdgcn_readfirstlane((int)__ockl_wfred_max_u32(key));
                    if ((key >> 8) + h + p >= k) {  // (key >> 8) = max score + 1 bounds the clique inside X
                        const uint32_t pv = 63u - (key & 0xFFu);
                        ++t;
                        const uint64_t bw = X & ~rdlane64(row, pv);
                        stkC = (lane == (uint32_t)t) ? X : stkC;
                        stkB = (lane == (uint32_t)t) ? bw : stkB;
                        const uint32_t M = h | (p << 8) | (pv << 20);
                        stkM = (lane == (uint32_t)t) ? M : stkM;
                    }
                }
            }
        }
        if (t < 0) break;
        // next branch of level t: non-pivot vertices first, the pivot last
        uint64_t Cc = rdlane64(stkC, (uint32_t)t), B = rdlane64(stkB, (uint32_t)t);
        const uint32_t M = (uint32_t)__builtin_amdgcn_readlane((int)stkM, t);
        const uint32_t hh = M & 0xFF, pp = (M >> 8) & 0xFFF, pv = M >> 20;
        const uint64_t nb = B & ~(1ull << pv);
        const bool last = nb == 0;
        const uint32_t b = last ? pv : (uint32_t)__builtin_ctzll(nb);
        X = Cc & rdlane64(row, b);
        h = hh + (last ? 0u : 1u);
        p = pp + (last ? 1u : 0u);
        if (last) {
            --t;  // the pivot branch replaces its level
        } else {
            Cc &= ~(1ull << b);
            B &= ~(1ull << b);
            stkC = (lane == (uint32_t)t) ? Cc : stkC;
            stkB = (lane == (uint32_t)t) ? B : stkB;
        }
        if (h <= k) pending = true;
    }
    if (lane == 0) nodes += my_nodes;
    return acc;
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
// so the SCT is only expanded while h < k-2.  h = k-3 closes with c_3 =
// triangles of G[X] (lane walks; with per-vertex credits for LOCAL), and the
// register kernels close h = k-4 with c_3 and c_4 of near-cliques from the
// complement.  A node is also dropped when
// h + p + max_j|X & row_j| + 1 < k (the largest clique inside X has at most
// maxIntersection + 1 vertices, cf. kclique/kckernels.cuh:2573-2584).
//
// LOCAL also credits every task vertex in lacc with the k-cliques of the
// subtree that contain it (the root with the task's count, by the caller).
// A subtree (H, P, X) with t = k - h holds the cliques H + S + Q, S a subset
// of P and Q a clique of G[X] with |S| + |Q| = t, so
//     held vertex  sum_j C(p, t-j) c_j(X)      pivot vertex  sum_j C(p-1, t-1-j) c_j(X)
//     x in X       sum_j C(p, t-j) c_{j-1}(X & row_x)
// which at a leaf (X empty) is the reference's hold/pivot split
// (kclique/kclique_local.cuh, kckernels_local.cuh).  The sets H and P of the
// node live beside X, and each stack level keeps its own copies.  Vertex sets
// H, P and X are disjoint and lane w owns the 64 vertices of word w, so the
// adds need no atomics; no add can wrap unless the task count does.
__device__ inline uint64_t binom_at(const CountArgs &a, int64_t n, int64_t r)
{
    return (r < 0 || n < 0 || r > n) ? 0ull : a.binom[(uint64_t)n * a.bstride + (uint64_t)r];
}

// Barrier for the state of one traversal: the whole block when a block is
// one wave, only the wave when the waves of a block run separate subtrees.
template <bool WIDE>
__device__ inline void sync_tree()
{
    if (WIDE) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    } else {
        __syncthreads();
    }
}

// The SCT from the root candidate set (all R task vertices, held h0, no
// pivots), or -- with init_x false -- from the candidate set already in X
// with h0 held and p0 pivot vertices (a subtree handed to one wave of a block).
template <int S>
__device__ uint32_t tri_lanes(const uint64_t *rows, uint32_t W, const uint64_t (&X)[S], uint32_t cx, uint32_t e2,
                              uint32_t *idx);
// CL4: the four-word kernel also closes h = k-4 near-cliques (the wide kernel
// has the registers for it; the one-wave four-word kernel keeps its budget)
template <int S, bool CL4 = false>
__device__ uint64_t pivot_reg(const CountArgs &a, uint32_t k, const uint64_t *rows, uint32_t R, uint32_t h0,
                              uint32_t *idx, uint64_t &nodes, uint32_t &ovf, uint32_t p0, uint64_t *crows);

// triangles of G[X] with X in LDS (W words): the lane walks with the words in registers
template <int S>
__device__ uint32_t tri_lds(const uint64_t *rows, uint32_t W, const uint64_t *X, uint32_t cx, uint32_t e2,
                            uint32_t *idx)
{
    uint64_t Xr[S];
#pragma unroll
    for (int w = 0; w < S; ++w) Xr[w] = (uint32_t)w < W ? uni64(X[w]) : 0ull;
    return tri_lanes<S>(rows, W, Xr, cx, e2, idx);
}

template <bool LOCAL, bool WIDE = false>
__device__ uint64_t pivot_count(const CountArgs &a, uint32_t P, uint32_t k, const uint64_t *rows, uint64_t *stk,
                                uint64_t *X, Scal *scal, uint32_t *cand, uint32_t R, uint32_t W, uint32_t h0,
                                uint64_t &nodes, uint32_t &ovf, uint64_t *lacc, uint32_t p0 = 0, bool init_x = true,
                                uint64_t *crows2 = nullptr, uint64_t *crows4 = nullptr)
{
    const int lane = lane_id();
    const uint64_t lt = lanemask_lt();
    const uint32_t wmax = a.wmax;
    const uint32_t NPp = 64 / P, part = lane / P, lip = lane % P;  // P lanes score one candidate
    const uint32_t lvl = LOCAL ? 4 * wmax : 2 * wmax;  // stack words per level
    uint64_t *Hc = X + wmax, *Pc = X + 2 * wmax;     // LOCAL: held / pivot sets of the node
    uint64_t acc = 0;
    // root candidate set = all R list vertices
    if (init_x)
        for (uint32_t w = lane; w < W; w += 64) {
            uint32_t lo = w * 64;
            X[w] = (R - lo >= 64) ? ~0ull : ((1ull << (R - lo)) - 1);
            if (LOCAL) Hc[w] = Pc[w] = 0;
        }
    sync_tree<WIDE>();
    int t = -1;
    bool pending = true;
    uint32_t h = h0, p = p0;
    uint64_t my_nodes = 0;
    // LOCAL: add vh to each held and vp to each pivot vertex of the node
    auto credit_hp = [&](uint64_t vh, uint64_t vp) {
        if (!LOCAL || (!vh && !vp)) return;
        for (uint32_t w = lane; w < W; w += 64) {
            if (vh)
                for (uint64_t x = Hc[w]; x; x &= x - 1) lacc[w * 64 + (uint32_t)__builtin_ctzll(x)] += vh;
            if (vp)
                for (uint64_t x = Pc[w]; x; x &= x - 1) lacc[w * 64 + (uint32_t)__builtin_ctzll(x)] += vp;
        }
    };
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
            sync_tree<WIDE>();
            if (cx == 0) {
                if (k >= h && k - h <= p) {
                    uint64_t idx = (uint64_t)p * a.bstride + (k - h);
                    if (a.binom_ovf[idx]) ovf = 1;
                    add_checked(acc, a.binom[idx], ovf);
                    if (LOCAL) credit_hp(a.binom[idx], binom_at(a, (int64_t)p - 1, (int64_t)(k - h) - 1));
                }
            } else if (h >= k) {
                add_checked(acc, 1, ovf);
                if (LOCAL) credit_hp(1, 0);
            } else if (h + 1 == k) {
                add_checked(acc, (uint64_t)p + cx, ovf);
                if (LOCAL) {
                    credit_hp((uint64_t)p + cx, 1);
                    for (uint32_t c = lane; c < cx; c += 64) lacc[cand[c]] += 1;
                }
            } else if (!LOCAL && cx <= 64 && h + 3 <= k && h + p + cx >= k) {
                // at most 64 candidates left: re-encode the subtree one word per
                // row (row j of the subproblem = ballot of bit cand[j] over the
                // rows of the candidates) and finish it in sct64
                const uint32_t mine = lane < cx ? cand[lane] : 0u;
                uint64_t row1 = 0;
                for (uint32_t j = 0; j < cx; ++j) {
                    const uint32_t xj = cand[j];
                    const uint64_t wv = lane < cx ? rows[(uint64_t)mine * W + (xj >> 6)] : 0ull;
                    const uint64_t col = __ballot((wv >> (xj & 63)) & 1ull);
                    row1 = (lane == j) ? col : row1;
                }
                add_checked(acc, sct64(a, k, row1, cx >= 64 ? ~0ull : ((1ull << cx) - 1), h, p, nodes, ovf), ovf);
                --my_nodes;  // counted again as the subtree's root
            } else if (!LOCAL && crows4 && cx > 128 && cx <= 256 && h + 3 <= k && h + p + cx >= k) {
                // 129..256 candidates: re-encode three or four words per row the
                // same way and finish the subtree in the four-word register SCT
                // (which closes an h = k-4 near-clique at its root)
                const uint32_t Wc = (cx + 63) / 64;
                uint32_t mt[4];
#pragma unroll
                for (int t = 0; t < 4; ++t) mt[t] = 64u * t + lane < cx ? cand[64u * t + lane] : 0u;
                for (uint32_t j = 0; j < cx; ++j) {
                    const uint32_t xj = cand[j];
                    uint64_t mine = 0;
#pragma unroll
                    for (int t = 0; t < 4; ++t) {
                        const uint64_t wv = 64u * t + lane < cx ? rows[(uint64_t)mt[t] * W + (xj >> 6)] : 0ull;
                        const uint64_t col = __ballot((wv >> (xj & 63)) & 1ull);
                        mine = lane == (uint32_t)t ? col : mine;
                    }
                    if (lane < Wc) crows4[(uint64_t)Wc * j + lane] = mine;
                }
                sync_tree<WIDE>();
                add_checked(acc, pivot_reg<4, true>(a, k, crows4, cx, h, cand, nodes, ovf, p, crows2), ovf);
                --my_nodes;
            } else if (!LOCAL && crows2 && cx <= 128 && h + 3 <= k && h + p + cx >= k) {
                // 65..128 candidates: re-encode two words per row (word t of new
                // row j = ballot of bit cand[j] over the rows of candidates
                // 64t + lane) and finish the subtree in the two-word register SCT
                const uint32_t m0 = cand[lane], m1 = lane + 64 < cx ? cand[lane + 64] : 0u;
                for (uint32_t j = 0; j < cx; ++j) {
                    const uint32_t xj = cand[j];
                    const uint64_t w0 = rows[(uint64_t)m0 * W + (xj >> 6)];
                    const uint64_t w1 = lane + 64 < cx ? rows[(uint64_t)m1 * W + (xj >> 6)] : 0ull;
                    const uint64_t c0 = __ballot((w0 >> (xj & 63)) & 1ull), c1 = __ballot((w1 >> (xj & 63)) & 1ull);
                    if (lane < 2) crows2[2 * j + lane] = lane ? c1 : c0;
                }
                sync_tree<WIDE>();
                add_checked(acc, pivot_reg<2>(a, k, crows2, cx, h, cand, nodes, ovf, p, nullptr), ovf);
                --my_nodes;
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
                    // h = k-2: x in X lies in C(p,1) + |X & row_x| of the subtree's cliques
                    if (LOCAL && h + 2 == k && lip == 0) lacc[j] += (uint64_t)p + sc;
                }
                uint64_t key = (lip == 0 && best_j != INF32) ? (((uint64_t)best_s << 32) | (INF32 - best_j)) : 0ull;
                key = wave_max64(key);
                uint32_t dummy = 0;
                const uint64_t tot_s = wave_sum64_checked(lip == 0 ? sum_s : 0ull, dummy);
                const uint32_t max_s = (uint32_t)(key >> 32);
                if (h + 2 == k) {
                    const uint64_t before = acc;
                    add_checked(acc, (uint64_t)p * (p - (p ? 1 : 0)) / 2, ovf);
                    add_checked(acc, (uint64_t)p * cx, ovf);
                    add_checked(acc, tot_s / 2, ovf);
                    if (LOCAL) credit_hp(acc - before, (p ? (uint64_t)p - 1 : 0) + cx);
                } else if (LOCAL && h + 3 == k) {
                    // h = k-3 with per-vertex credits: x in X lies in
                    // C(p,2) + p deg_X(x) + t_X(x) of the subtree's cliques (t_X(x) =
                    // edges of G[X] inside N(x), walked by the lane that owns x),
                    // a pivot in C(p-1,2) + (p-1)|X| + |E(X)|, a held vertex in all
                    sync_tree<WIDE>();  // (lacc of the score pass)
                    const uint64_t pp = p, cp2 = pp * (pp - (pp ? 1 : 0)) / 2;
                    uint64_t t3 = 0;
                    for (uint32_t c = lane; c < cx; c += 64) {
                        const uint32_t x = cand[c];
                        const uint64_t *rx = rows + (uint64_t)x * W;
                        uint64_t deg = 0, t2 = 0;
                        for (uint32_t w = 0; w < W; ++w) {
                            const uint64_t yw = X[w] & rx[w];
                            deg += (uint64_t)__popcll(yw);
                            for (uint64_t bits = yw; bits; bits &= bits - 1) {
                                const uint64_t *ry = rows + (uint64_t)(w * 64 + (uint32_t)__builtin_ctzll(bits)) * W;
                                for (uint32_t v = 0; v < W; ++v) t2 += (uint64_t)__popcll(X[v] & rx[v] & ry[v]);
                            }
                        }
                        lacc[x] += cp2 + pp * deg + t2 / 2;
                        t3 += t2 / 2;
                    }
                    uint32_t dummy = 0;
                    const uint64_t tri = wave_sum64_checked(t3, dummy) / 3;
                    const uint64_t e = tot_s / 2;
                    const uint64_t before = acc;
                    add_checked(acc, close_k3(a, p, cx, e, tri, ovf), ovf);
                    const uint64_t q = pp ? pp - 1 : 0;
                    credit_hp(acc - before, p ? q * (q - (q ? 1 : 0)) / 2 + q * cx + e : 0);
                    sync_tree<WIDE>();
                } else if (!LOCAL && h + 3 == k && W <= 16) {
                    // h = k-3: closed form with the triangles of G[X] (lane walks)
                    const uint32_t e2 = (uint32_t)tot_s;
                    const uint32_t tri = W <= 4 ? tri_lds<4>(rows, W, X, cx, e2, cand)
                                         : W <= 8 ? tri_lds<8>(rows, W, X, cx, e2, cand)
                                                  : tri_lds<16>(rows, W, X, cx, e2, cand);
                    add_checked(acc, close_k3(a, p, cx, tot_s / 2, tri, ovf), ovf);
                } else if (h + p + max_s + 1 >= k) {
                    const uint32_t pv = INF32 - (uint32_t)(key & 0xFFFFFFFFull);
                    ++t;
                    uint64_t *Cc = stk + (uint64_t)t * lvl, *B = Cc + wmax;
                    const uint64_t *rp = rows + (uint64_t)pv * W;
                    for (uint32_t w = lane; w < W; w += 64) {
                        uint64_t x = X[w];
                        Cc[w] = x;
                        B[w] = x & ~rp[w];
                        if (LOCAL) {  // level layout: Cc, B, held, pivots (wmax words each)
                            Cc[2 * wmax + w] = Hc[w];
                            Cc[3 * wmax + w] = Pc[w];
                        }
                    }
                    if (lane == 0) scal[t] = Scal{h, p, pv, 0};
                    sync_tree<WIDE>();
                }
            }
        }
        if (t < 0) break;
        // next branch of level t: non-pivot vertices first, the pivot last
        uint64_t *Cc = stk + (uint64_t)t * lvl, *B = Cc + wmax;
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
            if (LOCAL) {
                const uint64_t bit = (w == (b >> 6)) ? (1ull << (b & 63)) : 0ull;
                Hc[w] = Cc[2 * wmax + w] | (b != sc.pivot ? bit : 0ull);
                Pc[w] = Cc[3 * wmax + w] | (b == sc.pivot ? bit : 0ull);
            }
            if (w == (b >> 6)) {
                Cc[w] = c & ~(1ull << (b & 63));
                B[w] &= ~(1ull << (b & 63));
            }
        }
        sync_tree<WIDE>();
        h = sc.held + (b != sc.pivot ? 1u : 0u);
        p = sc.piv + (b == sc.pivot ? 1u : 0u);
        if (last) --t;  // the pivot branch replaces its level
        if (h <= k) pending = true;
    }
    if (lane == 0) nodes += my_nodes;
    return acc;
}

// ---- step 3c: pivoting with the bitmap in registers (R <= 64*S) -------------
// Lane l holds rows l, l+64, ..., l+64(S-1) of the bitmap (in VGPRs for the
// one- and two-word kernels; the four-word level reads them from LDS), so a
// pivot score is S*S AND+popcount on the lane's own rows; the rows a branch
// needs whole are broadcast-read from the LDS copy left by the build.  The SCT
// state is uniform: candidate set X and the level variables live in scalar
// registers, and the stack is spread across the lanes (lane t keeps level t in
// VGPRs, read back with v_readlane).  Arg-max and sums are wave reductions.

// word w of row b from the LDS copy (uniform address: a broadcast read)
__device__ inline uint64_t lds_row_word(const uint64_t *rows, uint32_t W, uint32_t b, int w)
{
    if ((uint32_t)w >= W) return 0ull;
    const uint64_t v = rows[(uint64_t)b * W + w];
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}

// Compaction: once a node's candidate set X has at most 64 vertices, its
// subtree is re-encoded as a 64-vertex subproblem (vertex j = the j-th member
// of X) and finished by sct64.  Lane i loads the full row of the i-th member
// from the LDS copy; new row j is the ballot over lanes of bit x_j of those
// rows (the bitmap is symmetric).  Cost ~ 6 instructions per member.
template <int S>
__device__ uint64_t compact_and_count(const CountArgs &a, uint32_t k, const uint64_t *rows, uint32_t W,
                                      const uint64_t (&X)[S], uint32_t *idx, uint32_t h, uint32_t p,
                                      uint64_t &nodes, uint32_t &ovf)
{
    const uint32_t lane = (uint32_t)lane_id();
    const uint64_t lt = lanemask_lt();
    uint32_t before = 0, cx = 0;
#pragma unroll
    for (int s = 0; s < S; ++s) {
        if ((X[s] >> lane) & 1ull) idx[before + (uint32_t)__popcll(X[s] & lt)] = 64u * s + lane;
        before += (uint32_t)__popcll(X[s]);
    }
    cx = before;
    sync_tree<true>();
    uint64_t o[S];
    const uint32_t mine = lane < cx ? idx[lane] : 0u;
#pragma unroll
    for (int w = 0; w < S; ++w) o[w] = (lane < cx && (uint32_t)w < W) ? rows[(uint64_t)mine * W + w] : 0ull;
    sync_tree<true>();  // idx is reused by the next compaction
    uint64_t row1 = 0;
    uint32_t j = 0;
#pragma unroll
    for (int w = 0; w < S; ++w) {
        for (uint64_t bits = X[w]; bits; bits &= bits - 1, ++j) {
            const uint32_t bj = (uint32_t)__builtin_ctzll(bits);
            const uint64_t col = __ballot((o[w] >> bj) & 1ull);
            row1 = (lane == j) ? col : row1;
        }
    }
    const uint64_t X1 = cx >= 64 ? ~0ull : ((1ull << cx) - 1);
    return sct64(a, k, row1, X1, h, p, nodes, ovf);
}

// Triangles of G[X] for X of up to 64*S vertices (rows of stride W in LDS) by
// lane-parallel forward walks.  Lane l takes the members l, 127-l, 128+l,
// 255-l, ... of X (in index order; the pairs balance the walks) and, for each
// member x, walks its forward neighbours F = N(x) & X above x lowest first:
// popping j leaves in F exactly the forward neighbours above j, so
// |F & N(j)| counts the triangles x < j < y once.  When more than half the
// pairs of X are edges the walk runs in the complement H instead (F = the
// non-neighbours above x) and Goodman's identity
//     t(G) = C(n,3) - 1/2 sum_x d_x (n-1-d_x) - t(H)
// gives the count; either way a lane walks min(density, 1-density) * n
// neighbours per member.  e2 = 2|E(X)|.  idx takes the member list.
template <int S>
__device__ uint32_t tri_lanes(const uint64_t *rows, uint32_t W, const uint64_t (&X)[S], uint32_t cx, uint32_t e2,
                              uint32_t *idx)
{
    const uint32_t lane = (uint32_t)lane_id();
    const uint64_t lt = lanemask_lt();
    uint32_t before = 0;
#pragma unroll
    for (int s = 0; s < S; ++s) {
        if ((X[s] >> lane) & 1ull) idx[before + (uint32_t)__popcll(X[s] & lt)] = 64u * s + lane;
        before += (uint32_t)__popcll(X[s]);
    }
    sync_tree<true>();
    const bool comp = e2 > cx * (cx - 1) / 2;  // |E| > C(n,2)/2
    const uint64_t flip = comp ? ~0ull : 0ull;
    constexpr int NR = S > 2 ? S : 2;  // member slots per lane
    uint32_t val = 0, cnt = 0;
    int r = -1;
    uint64_t F[S];
#pragma unroll
    for (int w = 0; w < S; ++w) F[w] = 0;
    for (;;) {
        bool empty = true;
#pragma unroll
        for (int w = 0; w < S; ++w) empty = empty && F[w] == 0;
        while (empty && r < NR - 1) {  // the lane's next member
            ++r;
            const uint32_t m = (r & 1) ? (128u * (uint32_t)(r >> 1) + 127u - lane) : (128u * (uint32_t)(r >> 1) + lane);
            if (m >= cx) continue;
            const uint32_t x = idx[m], xw = x >> 6, xb = x & 63;
            uint32_t dg = 0;
#pragma unroll
            for (int w = 0; w < S; ++w) {
                const uint64_t rw = (uint32_t)w < W ? rows[(uint64_t)x * W + w] : 0ull;
                dg += (uint32_t)__popcll(X[w] & rw);
                const uint64_t above = (uint32_t)w > xw ? ~0ull : ((uint32_t)w < xw || xb == 63) ? 0ull : (~0ull << (xb + 1));
                F[w] = X[w] & (rw ^ flip) & above;
                empty = empty && F[w] == 0;
            }
            if (comp) val += dg * (cx - 1 - dg);
        }
        if (!__ballot(!empty)) break;
        if (!empty) {
            uint32_t j = 0;
            bool found = false;
#pragma unroll
            for (int w = 0; w < S; ++w) {
                if (!found && F[w]) {
                    j = 64u * w + (uint32_t)__builtin_ctzll(F[w]);
                    F[w] &= F[w] - 1;
                    found = true;
                }
            }
#pragma unroll
            for (int w = 0; w < S; ++w)
                if ((uint32_t)w < W) cnt += (uint32_t)__popcll(F[w] & (rows[(uint64_t)j * W + w] ^ flip));
        }
    }
    val += 2 * cnt;
    const uint32_t tot = (uint32_t)__builtin_amdgcn_readfirstlane((int)__ockl_wfred_add_u32(val));
    sync_tree<true>();  // idx is reused
    return comp ? cx * (cx - 1) * (cx - 2) / 6 - tot / 2 : tot / 2;
}

// k34_complement for X of up to 64*S vertices (rows of stride W in LDS): lane
// l takes the vertices l + 64s, s < S, with their complement rows in S words
// each; the same uniform walk over u in X, the same four running sums per
// vertex and the same K4 walk over the vertices in three or more complement
// triangles, with the lane shares summed in 64 bits (C(256,4) * 24 does not
// fit 32).  Used at h = k-4 when X is a near-clique (the complement walks
// stay short).
template <int S>
__device__ CliqueCounts k34_complement_rows(const uint64_t *rows, uint32_t W, const uint64_t (&X)[S], uint32_t cx)
{
    const uint32_t lane = (uint32_t)lane_id();
    const int64_t n = cx;
    uint64_t Hl[S][S];
    int64_t d[S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const uint32_t v = lane + 64u * s;
        const bool mem = (X[s] >> lane) & 1ull;
        d[s] = 0;
#pragma unroll
        for (int w = 0; w < S; ++w) {
            const uint64_t rw = (mem && (uint32_t)w < W) ? rows[(uint64_t)v * W + w] : ~0ull;
            Hl[s][w] = X[w] & ~rw & (w == s ? ~(1ull << lane) : ~0ull);
            d[s] += __popcll(Hl[s][w]);
        }
    }
    uint32_t q[S], qE[S], t2[S], sd[S];
#pragma unroll
    for (int s = 0; s < S; ++s) q[s] = qE[s] = t2[s] = sd[s] = 0;
#pragma unroll
    for (int wu = 0; wu < S; ++wu) {
        for (uint64_t bits = X[wu]; bits; bits &= bits - 1) {
            const uint32_t u = 64u * wu + (uint32_t)__builtin_ctzll(bits);
            uint64_t Hu[S];
            uint32_t du = 0;
#pragma unroll
            for (int w = 0; w < S; ++w) {
                Hu[w] = X[w] & ~lds_row_word(rows, W, u, w) & (w == wu ? ~(1ull << (u & 63)) : ~0ull);
                du += (uint32_t)__popcll(Hu[w]);
            }
#pragma unroll
            for (int s = 0; s < S; ++s) {
                uint32_t c = 0;
#pragma unroll
                for (int w = 0; w < S; ++w) c += (uint32_t)__popcll(Hl[s][w] & Hu[w]);
                const uint32_t c2 = c * c, in = (uint32_t)(Hu[s] >> lane) & 1u;
                q[s] += c2;
                qE[s] += in * c2;
                t2[s] += in * c;
                sd[s] += in * du;
            }
        }
    }
    // K4 of H among the vertices in >= 3 H-triangles (bit l of word s = vertex l + 64s)
    uint64_t T3[S];
    uint32_t nt3 = 0;
#pragma unroll
    for (int s = 0; s < S; ++s) {
        T3[s] = __ballot(t2[s] >= 6);
        nt3 += (uint32_t)__popcll(T3[s]);
    }
    int64_t k4acc[S];
#pragma unroll
    for (int s = 0; s < S; ++s) k4acc[s] = 0;
    if (nt3 >= 4) {
#pragma unroll
        for (int wu = 0; wu < S; ++wu) {
            for (uint64_t ub = T3[wu]; ub; ub &= ub - 1) {
                const uint32_t u = 64u * wu + (uint32_t)__builtin_ctzll(ub);
                uint64_t Hu[S];
#pragma unroll
                for (int w = 0; w < S; ++w) {
                    const uint64_t above = w > wu ? ~0ull : w < wu ? 0ull : ((u & 63) == 63 ? 0ull : (~0ull << ((u & 63) + 1)));
                    Hu[w] = T3[w] & ~lds_row_word(rows, W, u, w) & above;
                }
#pragma unroll
                for (int wv = 0; wv < S; ++wv) {
                    for (uint64_t vb = Hu[wv]; vb; vb &= vb - 1) {
                        const uint32_t v = 64u * wv + (uint32_t)__builtin_ctzll(vb);
                        uint64_t Sv[S];
                        uint32_t ns = 0;
#pragma unroll
                        for (int w = 0; w < S; ++w) {
                            const uint64_t above = w > wv ? ~0ull : w < wv ? 0ull : ((v & 63) == 63 ? 0ull : (~0ull << ((v & 63) + 1)));
                            Sv[w] = Hu[w] & ~lds_row_word(rows, W, v, w) & above;
                            ns += (uint32_t)__popcll(Sv[w]);
                        }
                        if (ns < 2) continue;
#pragma unroll
                        for (int s = 0; s < S; ++s) {
                            if (!((Sv[s] >> lane) & 1ull)) continue;
                            uint32_t z = 0;
#pragma unroll
                            for (int w = 0; w < S; ++w) {
                                const uint64_t above = w > s ? ~0ull : w < s ? 0ull : (lane == 63 ? 0ull : (~0ull << (lane + 1)));
                                z += (uint32_t)__popcll(Sv[w] & Hl[s][w] & above);
                            }
                            k4acc[s] += z;
                        }
                    }
                }
            }
        }
    }
    int64_t c4l = 0, c3l = 0, sdl = 0;
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const int64_t dd = d[s], t = t2[s] / 2, iq = q[s], iqE = qE[s], isd = sd[s], it2 = t2[s];
        c4l += 12 * (n - 4) * dd * (dd - 1) - 8 * (n - 3) * t - 12 * (dd - 1) * (isd - dd) + 24 * t -
               4 * dd * (dd - 1) * (dd - 2) + 3 * ((iq - dd * dd) - (isd - dd)) + 24 * t * (dd - 2) - 6 * (iqE - it2) +
               24 * k4acc[s];
        c3l += -3 * dd * (n - 1 - dd) - 2 * t;
        sdl += dd;
    }
    uint32_t dummy = 0;
    const int64_t s_d = (int64_t)uni64(wave_sum64_checked((uint64_t)sdl, dummy));
    const int64_t s_4 = (int64_t)uni64(wave_sum64_checked((uint64_t)c4l, dummy));
    const int64_t s_3 = (int64_t)uni64(wave_sum64_checked((uint64_t)c3l, dummy));
    const int64_t m = s_d / 2;
    const int64_t c4_24 = n * (n - 1) * (n - 2) * (n - 3) - 12 * m * (n - 2) * (n - 3) + 12 * m * (m - 1) + s_4;
    const int64_t c3_6 = n * (n - 1) * (n - 2) + s_3;
    CliqueCounts r;
    r.e = (uint32_t)(n * (n - 1) / 2 - m);
    r.c3 = (uint32_t)(c3_6 / 6);
    r.c4 = (uint32_t)(c4_24 / 24);
    return r;
}

// Compaction to two words: once a node of the four-word kernel has 65..128
// candidates, its subtree is re-encoded as a subproblem of at most 128
// vertices (vertex j = the j-th member of X) with two-word rows in crows and
// finished by the two-word register SCT, whose scores and closed forms cost a
// quarter of the four-word ones.  Lanes l and l+64 load the full rows of
// members l and l+64; word t of new row j is the ballot over lanes of bit x_j
// of the rows of members 64t + lane (the bitmap is symmetric).
template <int S, bool CL4>
__device__ uint64_t pivot_reg(const CountArgs &a, uint32_t k, const uint64_t *rows, uint32_t R, uint32_t h0,
                              uint32_t *idx, uint64_t &nodes, uint32_t &ovf, uint32_t p0, uint64_t *crows);

template <int S>
__device__ uint64_t compact2_and_count(const CountArgs &a, uint32_t k, const uint64_t *rows, uint32_t W,
                                       const uint64_t (&X)[S], uint32_t *idx, uint64_t *crows, uint32_t h, uint32_t p,
                                       uint64_t &nodes, uint32_t &ovf)
{
    const uint32_t lane = (uint32_t)lane_id();
    const uint64_t lt = lanemask_lt();
    uint32_t cx = 0;
#pragma unroll
    for (int s = 0; s < S; ++s) {
        if ((X[s] >> lane) & 1ull) idx[cx + (uint32_t)__popcll(X[s] & lt)] = 64u * s + lane;
        cx += (uint32_t)__popcll(X[s]);
    }
    sync_tree<true>();
    const bool hi_ok = lane + 64 < cx;
    const uint32_t m0 = idx[lane], m1 = hi_ok ? idx[lane + 64] : 0u;  // cx > 64: every lane has a first member
    uint64_t o0[S], o1[S];
#pragma unroll
    for (int w = 0; w < S; ++w) {
        o0[w] = (uint32_t)w < W ? rows[(uint64_t)m0 * W + w] : 0ull;
        o1[w] = (hi_ok && (uint32_t)w < W) ? rows[(uint64_t)m1 * W + w] : 0ull;
    }
    sync_tree<true>();  // idx is reused by the two-word kernel
    uint32_t j = 0;
#pragma unroll
    for (int w = 0; w < S; ++w) {
        for (uint64_t bits = X[w]; bits; bits &= bits - 1, ++j) {
            const uint32_t bj = (uint32_t)__builtin_ctzll(bits);
            const uint64_t c0 = __ballot((o0[w] >> bj) & 1ull), c1 = __ballot((o1[w] >> bj) & 1ull);
            if (lane < 2) crows[2 * j + lane] = lane ? c1 : c0;
        }
    }
    sync_tree<true>();
    return pivot_reg<2>(a, k, crows, cx, h, idx, nodes, ovf, p, nullptr);
}

template <int S, bool CL4>
__device__ uint64_t pivot_reg(const CountArgs &a, uint32_t k, const uint64_t *rows, uint32_t R, uint32_t h0,
                              uint32_t *idx, uint64_t &nodes, uint32_t &ovf, uint32_t p0, uint64_t *crows)
{
    const uint32_t lane = (uint32_t)lane_id();
    k = uni32(k);
    R = uni32(R);
    h0 = uni32(h0);
    p0 = uni32(p0);
    const uint32_t W = (R + 63) / 64;
    if (S == 1) {
        const uint64_t r1 = lane < R ? rows[lane] : 0ull;
        return sct64(a, k, r1, R >= 64 ? ~0ull : ((1ull << R) - 1), h0, p0, nodes, ovf);
    }
    // rows of the lane in VGPRs for the two-word kernel; the four-word level
    // (a small share of the work once subtrees of <= 128 candidates are
    // compacted) reads them from LDS, which keeps its register budget at the
    // two-word kernel's
    constexpr bool REG = S <= 2;
    constexpr int SR = REG ? S : 1;
    uint64_t row[SR][SR];
#pragma unroll
    for (int s = 0; s < SR; ++s) {
        const uint32_t j = lane + 64u * s;
#pragma unroll
        for (int w = 0; w < SR; ++w) row[s][w] = (REG && j < R && (uint32_t)w < W) ? rows[(uint64_t)j * W + w] : 0ull;
    }
    auto row_word = [&](int s, int w) -> uint64_t {
        if constexpr (REG) {
            return row[s][w];
        } else {
            const uint32_t j = lane + 64u * s;
            return (j < R && (uint32_t)w < W) ? rows[(uint64_t)j * W + w] : 0ull;
        }
    };
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
    uint32_t h = h0, p = p0;
    uint64_t acc = 0, my_nodes = 0;
    for (;;) {
        if (pending) {
            pending = false;
            uint32_t cx = 0;
#pragma unroll
            for (int w = 0; w < S; ++w) cx += (uint32_t)__popcll(X[w]);
            if (S == 4 && cx > 64 && cx <= 128 && h + 3 <= k && h + p + cx >= k) {
                // a subtree of at most 128 candidates: two words per row from here on
                add_checked(acc, compact2_and_count<S>(a, k, rows, W, X, idx, crows, h, p, nodes, ovf), ovf);
            } else if (cx <= 64 && cx > 0 && h + 3 <= k && h + p + cx >= k) {
                // a subtree of at most 64 candidates: one word per row from here on
                add_checked(acc, compact_and_count<S>(a, k, rows, W, X, idx, h, p, nodes, ovf), ovf);
            } else if ((++my_nodes, cx == 0)) {
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
                // scores |X & row_j| of the lane's member rows; key packs
                // (score + 1, 0xFFFF - j), so the wave max names the smallest j of
                // the largest score
                uint32_t key = 0, lsum = 0;
#pragma unroll
                for (int s = 0; s < S; ++s) {
                    uint32_t v = 0;
#pragma unroll
                    for (int w = 0; w < S; ++w) v += (uint32_t)__popcll(X[w] & row_word(s, w));
                    if ((X[s] >> lane) & 1ull) {
                        lsum += v;
                        const uint32_t kk = ((v + 1) << 16) | (0xFFFFu - (64u * s + lane));
                        key = kk > key ? kk : key;
                    }
                }
                uint32_t e2 = 0;
                if (h + 4 >= k) e2 = (uint32_t)__builtin_amdgcn_readfirstlane((int)__ockl_wfred_add_u32(lsum));
                if ((S == 2 || (S == 4 && CL4)) && h + 4 == k && 4 * (uint64_t)e2 >= 3ull * cx * (cx - 1)) {
                    // h = k-4 over a near-clique (complement density <= 1/4): closed form
                    const CliqueCounts cc = k34_complement_rows<S>(rows, W, X, cx);
                    add_checked(acc, binom_chk(a, p, 4, ovf), ovf);
                    add_checked(acc, mul_chk(binom_chk(a, p, 3, ovf), cx, ovf), ovf);
                    add_checked(acc, mul_chk(binom_chk(a, p, 2, ovf), cc.e, ovf), ovf);
                    add_checked(acc, mul_chk(p, cc.c3, ovf), ovf);
                    add_checked(acc, cc.c4, ovf);
                } else if (h + 3 >= k) {
                    // (e2 = twice the edges of G[X])
                    const uint64_t pp = p;
                    if (h + 2 == k) {
                        add_checked(acc, pp * (pp - (pp ? 1 : 0)) / 2, ovf);
                        add_checked(acc, pp * cx, ovf);
                        add_checked(acc, e2 / 2, ovf);
                    } else {
                        // h = k-3: C(p,3) + C(p,2)|X| + p|E(X)| + triangles(X)
                        const uint32_t tri = tri_lanes<S>(rows, W, X, cx, e2, idx);
                        add_checked(acc, pp * (pp - (pp ? 1 : 0)) * (pp >= 2 ? pp - 2 : 0) / 6, ovf);
                        add_checked(acc, pp * (pp - (pp ? 1 : 0)) / 2 * cx, ovf);
                        add_checked(acc, pp * (e2 / 2), ovf);
                        add_checked(acc, tri, ovf);
                    }
                } else if ((((key = (uint32_t)__builtin_amdgcn_readfirstlane((int)__ockl_wfred_max_u32(key))) >> 16) +
                            h + p) >= k) {  // (key >> 16) = max score + 1 bounds the clique inside X
                    const uint32_t pv = 0xFFFFu - (key & 0xFFFFu);
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
__device__ __attribute__((always_inline)) uint32_t stage_task(const Cur &g, uint32_t task, uint32_t *list,
                                                              uint64_t &bytes)
{
    const int lane = lane_id();
    if (!g.gran) {
        const uint32_t b = g.rowptr[task], e = g.rowptr[task + 1];
        const uint32_t R = e - b;
        bytes += 8 + 4ull * R;
        for (uint32_t i = lane; i < R; i += 64) list[i] = g.col[b + i];
        return R;
    }
    const uint32_t u = g.src[task], v = g.col[task];
    const uint32_t ub = g.rowptr[u], ue = g.rowptr[u + 1], vb = g.rowptr[v], ve = g.rowptr[v + 1];
    const uint32_t du = ue - ub, dv = ve - vb;
    bytes += 24 + 4ull * (du + dv);
    const uint32_t *sa = g.col + (du <= dv ? ub : vb), *sb = g.col + (du <= dv ? vb : ub);
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

// Per-wave bookkeeping shared by the kernels: the graph of the current task,
// the running count of the current query (flushed to qcount[] when the query
// changes and at the end), and the launch statistics.
struct WaveAcc {
    uint32_t q = INF32;
    uint64_t run = 0;
    uint32_t run_ovf = 0;
    uint64_t bytes = 0, ntask = 0, nodes = 0, tmax = 0, tsum = 0;
    uint32_t any_ovf = 0;
};

__device__ __attribute__((always_inline)) void flush_query(const CountArgs &a, WaveAcc &w)
{
    if (w.q == INF32) return;
    if (lane_id() == 0) {
        unsigned long long old = atomicAdd((unsigned long long *)&a.qcount[w.q], (unsigned long long)w.run);
        if (old + w.run < old || w.run_ovf) {
            atomicOr(&a.qovf[w.q], 1u);
        }
    }
    w.any_ovf |= w.run_ovf;
    w.run = 0;
    w.run_ovf = 0;
}

__device__ __attribute__((always_inline)) Cur select_task(const CountArgs &a, uint64_t code, WaveAcc &w,
                                                          uint32_t &task)
{
    const uint32_t q = (uint32_t)(code >> 32);
    task = (uint32_t)code;
    if (q != w.q) {
        flush_query(a, w);
        w.q = q;
    }
    const QDesc d = a.qd[q];
    return Cur{d.rowptr, d.col, d.src, d.k, d.gran, d.psize};
}

__device__ __attribute__((always_inline)) void finish_wave(const CountArgs &a, WaveAcc &w)
{
    flush_query(a, w);
    uint32_t nd_wrap = 0;
    const uint64_t nd = wave_sum64_checked(w.nodes, nd_wrap);  // tree nodes are per-lane shares
    if (lane_id() == 0) {
        atomicAdd((unsigned long long *)&a.acc[1], (unsigned long long)w.bytes);
        atomicAdd((unsigned long long *)&a.acc[2], (unsigned long long)w.ntask);
        atomicAdd((unsigned long long *)&a.acc[3], (unsigned long long)nd);
        atomicMax((unsigned long long *)&a.acc[4], (unsigned long long)w.tmax);
        atomicAdd((unsigned long long *)&a.acc[5], (unsigned long long)w.tsum);
    }
}

// register-bitmap pivoting kernel for tasks of at most 64*S vertices
// waves per SIMD the register budget is held to: 4 for S = 4 (<= 128 VGPRs), 6 for S = 2, 8 for S = 1
template <int S>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(S == 4 ? 4 : S == 2 ? 6 : 8)))
void k_pivot_reg(CountArgs a)
{
    extern __shared__ uint64_t lds[];
    const int lane = lane_id();
    const Layout L = make_layout(a.rmax, a.wmax, 0, 0, true);
    uint32_t *list = reinterpret_cast<uint32_t *>(lds + L.list);
    uint32_t *offs = reinterpret_cast<uint32_t *>(lds + L.offs);
    uint32_t *starts = reinterpret_cast<uint32_t *>(lds + L.starts);
    uint64_t *rows = lds + L.rows;
    // (S = 4) the two-word rows of a compacted subtree, 128 x 2 words, over
    // list and starts (256 u32 each, dead once the bitmap is built)
    uint64_t *crows = lds + L.list;
    WaveAcc w;
    for (;;) {
        uint32_t start = 0;
        if (lane == 0) start = atomicAdd(a.qhead, a.chunk);
        start = uni32(start);
        if (start >= a.ntasks) break;
        const uint32_t stop = min(a.ntasks, start + a.chunk);
        for (uint32_t ti = start; ti < stop; ++ti) {
            const uint64_t t0 = wall_clock64();
            uint32_t task;
            const Cur g = select_task(a, a.tasks[ti], w, task);
            ++w.ntask;
            const uint32_t R = stage_task(g, task, list, w.bytes);
            if (g.gran && g.k == 3) {  // an edge task at k=3 adds |T| (no bitmap needed)
                add_checked(w.run, R, w.run_ovf);
                __syncthreads();
                continue;
            }
            __syncthreads();
            build_rows<true>(g, list, offs, starts, rows, R, (R + 63) / 64, w.bytes);
            add_checked(w.run, pivot_reg<S>(a, g.k, rows, R, g.gran ? 2u : 1u, offs, w.nodes, w.run_ovf, 0u, crows),
                        w.run_ovf);
            __syncthreads();
            const uint64_t dt = wall_clock64() - t0;
            w.tmax = dt > w.tmax ? dt : w.tmax;
            w.tsum += dt;
        }
    }
    finish_wave(a, w);
}

// ---- open queues: the serving loop's resident launches ------------------------
// The server keeps resident launches of k_pivot_open while queries arrive,
// each draining one FIFO queue of tasks.  A query is a segment (base, slot,
// task count) of the queue's task index space.  The host admits a query with
// plain stores to pinned host memory (the slot's descriptor, the segment,
// then the queue's task total with release), so admission needs no HIP call
// and no free CU while the launch runs.  A wave takes the next task index
// with one fetch-add on a device word; an index beyond what the device knows
// is published is held until it is (at most every POLL_TICKS one wave reads
// the host's total and raises the device's copy).  The host ends a launch by
// publishing a stop segment of one index per wave, so no index is ever
// handed back; a wave also ends by itself when it has held an index for
// safety_ticks (the host is gone).  A task's size is computed by the wave
// that stages it, and tasks below the clique part still to find add 0.  A
// wave keeps the running count of its slot and flushes it, with the number
// of tasks it finished, when the slot changes and before it waits; the flush
// that completes the slot's task count resets the slot's sums and writes the
// result to the slot's mailbox in host memory.
constexpr uint32_t OPEN_STOP = 0xFFFFFFFFu;  // slot of a stop segment

struct OpenArgs {
    CountArgs c;              // binom, qcount / qovf per slot, acc (statistics)
    uint64_t *counter;        // next task index
    uint64_t *known;          // task indices the device knows are published
    uint64_t *last_poll;      // wall clock of the last read of the host's total
    uint64_t *qdone;          // per slot: tasks finished
    const uint64_t *htotal;   // host: task indices published
    const uint64_t *hnseg;    // host: segments published
    const uint64_t *hseg;     // host: per segment base, then (slot << 32) | task count
    const OpenSlot *hslot;    // host: per slot descriptor
    OpenMail *mail;           // host: per slot result
    uint64_t safety_ticks;    // wall-clock ticks (100 MHz) a wave waits for its index at most
};

constexpr uint64_t POLL_TICKS = 200;  // 2 us

__device__ inline uint64_t ld_agent(uint64_t *p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
__device__ inline uint64_t ld_host(const uint64_t *p)
{
    return __hip_atomic_load(const_cast<uint64_t *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// lane 0: read the host's task total, at most once per POLL_TICKS over the launch
__device__ inline void open_refresh(const OpenArgs &oa)
{
    const uint64_t now = wall_clock64();
    uint64_t last = ld_agent(oa.last_poll);
    if (now - last < POLL_TICKS) return;
    if (!__hip_atomic_compare_exchange_strong(oa.last_poll, &last, now, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                              __HIP_MEMORY_SCOPE_AGENT))
        return;
    const uint64_t h = ld_host(oa.htotal);
    if (h > ld_agent(oa.known)) __hip_atomic_fetch_max(oa.known, h, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// the segment a wave last resolved an index in (indices a wave takes only grow)
struct SegCache {
    uint64_t seg = 0, base = 0, desc = 0;
    bool valid = false;
};

// lane 0: the segment holding published index i (segments are in base
// order; binary search in host memory from the cached one)
__device__ inline void open_resolve(const OpenArgs &oa, SegCache &sc, uint64_t i)
{
    if (sc.valid && i >= sc.base && i - sc.base < (uint32_t)sc.desc) return;
    uint64_t lo = sc.valid ? sc.seg : 0, hi = ld_host(oa.hnseg);  // base[lo] <= i < total
    while (hi - lo > 1) {
        const uint64_t mid = (lo + hi) >> 1;
        if (ld_host(oa.hseg + 2 * mid) <= i) lo = mid;
        else hi = mid;
    }
    sc.seg = lo;
    sc.base = ld_host(oa.hseg + 2 * lo);
    sc.desc = ld_host(oa.hseg + 2 * lo + 1);
    sc.valid = true;
}

// flush the running count of the current slot and the tasks finished for it;
// the flush that completes the slot publishes its result and resets its sums
__device__ __attribute__((always_inline)) void open_flush(const OpenArgs &oa, WaveAcc &w, uint32_t &pend,
                                                         uint64_t total, uint32_t seq)
{
    if (w.q == INF32) return;
    const CountArgs &a = oa.c;
    if (lane_id() == 0 && pend) {
        const uint32_t q = w.q;
        const unsigned long long old = atomicAdd((unsigned long long *)&a.qcount[q], (unsigned long long)w.run);
        if (old + w.run < old || w.run_ovf) atomicOr(&a.qovf[q], 1u);
        const uint64_t before = __hip_atomic_fetch_add(&oa.qdone[q], (uint64_t)pend, __ATOMIC_ACQ_REL,
                                                       __HIP_MEMORY_SCOPE_AGENT);
        if (before + pend == total) {
            const uint64_t cnt = ld_agent(&a.qcount[q]);
            const uint32_t ovf = __hip_atomic_load(&a.qovf[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&a.qcount[q], 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&a.qovf[q], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&oa.qdone[q], 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&oa.mail[q].count, cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            __hip_atomic_store(&oa.mail[q].ovf, ovf, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            __hip_atomic_store(&oa.mail[q].seq, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
        }
    }
    w.any_ovf |= w.run_ovf;
    w.run = 0;
    w.run_ovf = 0;
    pend = 0;
    w.q = INF32;  // the slot may hold another query by the time this wave sees it again
}

// one-word rows: every task of the graphs the resident launches take fits
// 64 vertices, so the kernel keeps the one-word kernel's register budget
// (eight waves per SIMD)
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(8))) void k_pivot_open(OpenArgs oa)
{
    extern __shared__ uint64_t lds[];
    const CountArgs &a = oa.c;
    const int lane = lane_id();
    const Layout L = make_layout(a.rmax, a.wmax, 0, 0, true);
    uint32_t *list = reinterpret_cast<uint32_t *>(lds + L.list);
    uint32_t *offs = reinterpret_cast<uint32_t *>(lds + L.offs);
    uint32_t *starts = reinterpret_cast<uint32_t *>(lds + L.starts);
    uint64_t *rows = lds + L.rows;
    WaveAcc w;
    SegCache sc;
    uint32_t pend = 0, seq = 0;
    uint64_t total = 0;
    Cur g{};
    for (;;) {
        // lane 0 takes an index and waits until it is published
        uint64_t idx = 0;
        uint32_t ready = 0;
        if (lane == 0) {
            idx = __hip_atomic_fetch_add(oa.counter, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            open_refresh(oa);
            ready = idx < ld_agent(oa.known);
        }
        ready = uni32(__shfl(ready, 0, 64));
        if (!ready) {
            open_flush(oa, w, pend, total, seq);  // nothing held back while the wave waits
            if (lane == 0) {
                // back off while nothing is published: 1, 2, 4 .. 16 sleeps of
                // about 3.4 us between polls, so idle waves do not crowd the
                // device words (and the memory channel) the busy ones use
                const uint64_t t0 = wall_clock64();
                uint32_t naps = 1;
                while (!ready && wall_clock64() - t0 < oa.safety_ticks) {
                    for (uint32_t z = 0; z < naps; ++z) __builtin_amdgcn_s_sleep(127);
                    naps = naps < 16 ? 2 * naps : 16;
                    open_refresh(oa);
                    ready = idx < ld_agent(oa.known);
                }
            }
            ready = uni32(__shfl(ready, 0, 64));
            if (!ready) break;  // the host is gone: give up the index
        }
        uint32_t slot = 0, task = 0;
        if (lane == 0) {
            open_resolve(oa, sc, idx);
            slot = (uint32_t)(sc.desc >> 32);
            task = (uint32_t)(idx - sc.base);
        }
        slot = uni32(__shfl(slot, 0, 64));
        task = uni32(__shfl(task, 0, 64));
        if (slot == OPEN_STOP) break;
        if (slot != w.q) {
            open_flush(oa, w, pend, total, seq);
            w.q = slot;
            // the slot's descriptor, from host memory (lanes 0..5 read one word each)
            const uint64_t *hs = reinterpret_cast<const uint64_t *>(oa.hslot + slot);
            const uint64_t word = lane < (int)(sizeof(OpenSlot) / 8) ? ld_host(hs + lane) : 0ull;
            OpenSlot d;
            uint64_t *dw = reinterpret_cast<uint64_t *>(&d);
#pragma unroll
            for (int i = 0; i < (int)(sizeof(OpenSlot) / 8); ++i) dw[i] = uni64(rdlane64(word, (uint32_t)i));
            g = Cur{d.rowptr, d.col, d.src, d.k, d.gran, 1u};
            total = d.total;
            seq = d.seq;
        }
        ++w.ntask;
        ++pend;
        const uint32_t R = stage_task(g, task, list, w.bytes);
        const uint32_t need = g.gran ? g.k - 2 : g.k - 1;
        if (R >= need) {
            if (g.gran && g.k == 3) {
                add_checked(w.run, R, w.run_ovf);
            } else {
                __syncthreads();
                build_rows<true>(g, list, offs, starts, rows, R, (R + 63) / 64, w.bytes);
                add_checked(w.run,
                            pivot_reg<1>(a, g.k, rows, R, g.gran ? 2u : 1u, offs, w.nodes, w.run_ovf, 0u, nullptr),
                            w.run_ovf);
            }
        }
        __syncthreads();
    }
    open_flush(oa, w, pend, total, seq);
    finish_wave(a, w);
}

// ---- the generic kernel: LDS (or HBM) bitmap, orientation or pivoting ------
// LOCAL: per-vertex counts of one query into a.local (the -m lkc path): each
// task zeroes its vertex accumulators, the traversal fills them, and at the
// end one HBM atomic per credited task vertex plus one per root vertex (the
// root lies in every clique of its task) publishes them.
template <bool LOCAL>
__device__ __attribute__((always_inline)) void publish_local(const CountArgs &a, const Cur &g, uint32_t task,
                                                             const uint32_t *list, const uint64_t *lacc, uint32_t R,
                                                             uint64_t task_count)
{
    if (!LOCAL) return;
    const int lane = lane_id();
    for (uint32_t i = lane; i < R; i += 64)
        if (lacc[i]) atomicAdd((unsigned long long *)&a.local[list[i]], (unsigned long long)lacc[i]);
    if (lane == 0 && task_count) {
        const uint32_t u = g.gran ? g.src[task] : task;
        atomicAdd((unsigned long long *)&a.local[u], (unsigned long long)task_count);
        if (g.gran) atomicAdd((unsigned long long *)&a.local[g.col[task]], (unsigned long long)task_count);
    }
}

template <bool PIVOT, bool GLOBAL, bool LOCAL>
__global__ __launch_bounds__(64) void k_count(CountArgs a)
{
    extern __shared__ uint64_t lds[];
    const int lane = lane_id();
    uint64_t *ws = GLOBAL ? a.gscr + (uint64_t)blockIdx.x * a.gscr_words : lds;
    const Layout L = make_layout(a.rmax, a.wmax, a.levels, a.nparts, PIVOT, LOCAL);
    uint32_t *misc = reinterpret_cast<uint32_t *>(ws + L.misc);
    uint32_t *list = reinterpret_cast<uint32_t *>(ws + L.list);
    uint32_t *offs = reinterpret_cast<uint32_t *>(ws + L.offs);
    uint32_t *starts = reinterpret_cast<uint32_t *>(ws + L.starts);
    uint64_t *rows = ws + L.rows;
    uint64_t *stk = ws + L.stk;
    uint64_t *X = ws + L.x;
    Scal *scal = reinterpret_cast<Scal *>(ws + L.scal);
    uint64_t *lacc = ws + L.lacc;
    uint32_t *path = reinterpret_cast<uint32_t *>(ws + L.path);
    WaveAcc w;
    for (;;) {
        uint32_t start = 0;
        if (lane == 0) start = atomicAdd(a.qhead, a.chunk);
        start = uni32(start);
        if (start >= a.ntasks) break;
        const uint32_t stop = min(a.ntasks, start + a.chunk);
        for (uint32_t ti = start; ti < stop; ++ti) {
            const uint64_t t0 = wall_clock64();
            uint32_t task;
            const Cur g = select_task(a, a.tasks[ti], w, task);
            const uint32_t k = g.k;
            ++w.ntask;
            const uint32_t R = stage_task(g, task, list, w.bytes);
            if (LOCAL)
                for (uint32_t i = lane; i < R; i += 64) lacc[i] = 0;
            if (g.gran && k == 3) {  // an edge task at k=3 adds |T| (no bitmap needed)
                add_checked(w.run, R, w.run_ovf);
                if (LOCAL) {
                    for (uint32_t i = lane; i < R; i += 64) lacc[i] = 1;
                    __syncthreads();
                    publish_local<LOCAL>(a, g, task, list, lacc, R, R);
                    __syncthreads();
                }
                continue;
            }
            if (lane == 0) misc[0] = 0;
            __syncthreads();
            const uint32_t W = (R + 63) / 64;
            uint64_t task_count;
            if (!PIVOT) {
                // (k-1)-cliques of the DAG on N+(u) / (k-2)-cliques of the DAG on T
                build_rows<false>(g, list, offs, starts, rows, R, W, w.bytes);
                uint32_t povf = 0;
                uint64_t part = orient_count<LOCAL>(a, g.psize, rows, stk, misc, R, W, g.gran ? k - 2 : k - 1,
                                                    w.nodes, povf, lacc, path);
                task_count = wave_sum64_checked(part, povf);
                w.run_ovf |= (wave_min32(povf ? 0u : 1u) == 0) ? 1u : 0u;
                add_checked(w.run, task_count, w.run_ovf);
            } else {
                // SCT rooted at the node (held 1) or the edge (held 2)
                build_rows<true>(g, list, offs, starts, rows, R, W, w.bytes);
                task_count = pivot_count<LOCAL>(a, g.psize, k, rows, stk, X, scal, offs, R, W, g.gran ? 2u : 1u,
                                                w.nodes, w.run_ovf, lacc);
                add_checked(w.run, task_count, w.run_ovf);
            }
            __syncthreads();
            publish_local<LOCAL>(a, g, task, list, lacc, R, task_count);
            if (LOCAL) __syncthreads();
            const uint64_t dt = wall_clock64() - t0;
            w.tmax = dt > w.tmax ? dt : w.tmax;
            w.tsum += dt;
        }
    }
    finish_wave(a, w);
}

// ---- the wide kernel: one block of NW waves per large task -----------------
// Tasks of 257..1024 vertices (LDS) and above (HBM slice) hold a bitmap of up
// to 128 KB, so a one-wave block leaves most of a CU idle.  Here the NW waves
// of a block share one task: they build its bitmap together, then wave 0
// walks the SCT's pivot chain from the root (node j = the j-th pivot child:
// h0 held, j pivots) down to a node of at most 64 candidates or to the chain
// cap, and every non-pivot branch b hanging off chain node j -- candidates
// (X_j minus the non-pivot branches of j before b) & N(b), h0+1 held, j
// pivots -- plus the chain's last node become entries that the waves take
// from an LDS counter and finish with their own stacks
// (pivot_count<false, true>).  The chain is the part of the tree that one
// wave would otherwise walk alone: the cap is 32 levels when the chain lives
// in LDS and the whole chain (up to 1024 levels) in the HBM slice.  Blocks
// have up to 8 waves (512 threads), which leaves the register SCT that
// pivot_count hands subtrees of <= 256 candidates to its full register
// budget; each wave has its own LDS areas for those subtrees' rows.
struct WideLayout {
    uint64_t ctl, list, offs, starts, chain, rows, wave, per_wave, total;
};

__host__ __device__ inline WideLayout make_wide_layout(uint32_t rmax, uint32_t wmax, uint32_t levels, uint32_t nw,
                                                       uint32_t chain)
{
    WideLayout L;
    L.ctl = 0;
    L.list = L.ctl + 8;
    L.offs = L.list + (rmax + 1) / 2;
    L.starts = L.offs + (rmax + 2) / 2;
    L.chain = L.starts + (rmax + 1) / 2;         // per chain level: candidates, non-pivot branches
    L.rows = L.chain + (uint64_t)chain * 2 * wmax;
    L.wave = L.rows + (uint64_t)rmax * wmax;
    // + the rows of compacted subtrees: 128 x 2 and 256 x 4 words
    L.per_wave = (uint64_t)levels * 2 * wmax + wmax + (uint64_t)levels * 2 + (rmax + 1) / 2 + 256 + 1024;
    L.total = L.wave + L.per_wave * nw;
    return L;
}

// bitmap rows of the task by every thread of the block (build_rows with the
// row offsets scanned by wave 0)
__device__ void build_rows_block(const Cur &g, const uint32_t *list, uint32_t *offs, uint32_t *starts, uint64_t *rows,
                                 uint32_t R, uint32_t W, uint64_t &bytes)
{
    const uint32_t tid = threadIdx.x, nt = blockDim.x, lane = tid & 63;
    for (uint32_t i = tid; i < R * W; i += nt) rows[i] = 0;
    if (tid < 64) {
        uint32_t carry = 0;
        for (uint32_t base = 0; base < R; base += 64) {
            const uint32_t i = base + lane;
            uint32_t d = 0;
            if (i < R) {
                const uint32_t w = list[i];
                const uint32_t st = g.rowptr[w];
                d = g.rowptr[w + 1] - st;
                starts[i] = st;
            }
            const uint32_t inc = wave_incl_scan32(d);
            if (i < R) offs[i] = carry + inc - d;
            carry += __shfl(inc, 63, 64);
        }
        if (lane == 0) offs[R] = carry;
        bytes += 8ull * R + 4ull * carry;
    }
    __syncthreads();
    const uint32_t total = offs[R];
    const uint32_t lo_v = list[0], hi_v = list[R - 1];
    uint32_t lo = 0;  // row of entry f (f only grows)
    for (uint32_t f = tid; f < total; f += nt) {
        while (offs[lo + 1] <= f) ++lo;
        const uint32_t x = g.col[starts[lo] + (f - offs[lo])];
        if (x < lo_v || x > hi_v) continue;
        const uint32_t j = lower_bound_u32(list, R, x);
        if (j < R && list[j] == x) {
            atomicOr((unsigned long long *)&rows[(uint64_t)lo * W + (j >> 6)], 1ull << (j & 63));
            atomicOr((unsigned long long *)&rows[(uint64_t)j * W + (lo >> 6)], 1ull << (lo & 63));
        }
    }
    __syncthreads();
}

template <bool GLOBAL>
__global__ __launch_bounds__(512) void k_pivot_wide(CountArgs a)
{
    extern __shared__ uint64_t lds[];
    const uint32_t tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, NW = blockDim.x >> 6;
    uint64_t *ws = GLOBAL ? a.gscr + (uint64_t)blockIdx.x * a.gscr_words : lds;
    const WideLayout L = make_wide_layout(a.rmax, a.wmax, a.levels, NW, a.chain);
    uint32_t *ctl = reinterpret_cast<uint32_t *>(ws + L.ctl);  // task, R, next branch, branches, pivot
    uint32_t *list = reinterpret_cast<uint32_t *>(ws + L.list);
    uint32_t *offs = reinterpret_cast<uint32_t *>(ws + L.offs);  // row offsets, then the entries (j << 16 | b)
    uint32_t *starts = reinterpret_cast<uint32_t *>(ws + L.starts);
    uint64_t *rows = ws + L.rows;
    uint64_t *stk = ws + L.wave + (uint64_t)wv * L.per_wave;
    uint64_t *X = stk + (uint64_t)a.levels * 2 * a.wmax;
    Scal *scal = reinterpret_cast<Scal *>(X + a.wmax);
    uint32_t *cand = reinterpret_cast<uint32_t *>(X + a.wmax + (uint64_t)a.levels * 2);
    uint64_t *crows2 = X + a.wmax + (uint64_t)a.levels * 2 + (a.rmax + 1) / 2;  // 128 x 2 words
    uint64_t *crows4 = crows2 + 256;                                           // 256 x 4 words
    WaveAcc w;
    uint64_t no_bytes = 0;
    for (;;) {
        if (tid == 0) ctl[0] = atomicAdd(a.qhead, 1u);
        __syncthreads();
        const uint32_t ti = ctl[0];
        if (ti >= a.ntasks) break;
        const uint64_t t0 = wall_clock64();
        uint32_t task;
        const Cur g = select_task(a, a.tasks[ti], w, task);
        const uint32_t k = g.k, h0 = g.gran ? 2u : 1u;
        if (wv == 0) {
            const uint32_t r = stage_task(g, task, list, w.bytes);
            if (lane == 0) ctl[1] = r;
            ++w.ntask;
        }
        __syncthreads();
        const uint32_t R = ctl[1], W = (R + 63) / 64;
        if (g.gran && k == 3) {  // an edge task at k=3 adds |T| (no bitmap needed)
            if (wv == 0) add_checked(w.run, R, w.run_ovf);
            __syncthreads();
            continue;
        }
        build_rows_block(g, list, offs, starts, rows, R, W, wv == 0 ? w.bytes : no_bytes);
        // the pivot chain, by wave 0: closed forms at the root, or the entries
        uint64_t *chX = ws + L.chain, *chB = chX + a.wmax;  // level j at chX + 2 j wmax
        if (wv == 0) {
            uint32_t ne = 0, endj = INF32;
            for (uint32_t x = lane; x < W; x += 64) {
                const uint32_t lo = x * 64;
                chX[x] = (R - lo >= 64) ? ~0ull : ((1ull << (R - lo)) - 1);
            }
            sync_tree<true>();
            for (uint32_t j = 0;; ++j) {
                const uint64_t *Xj = chX + (uint64_t)j * 2 * a.wmax;
                uint64_t *Bj = chB + (uint64_t)j * 2 * a.wmax;
                const uint32_t p = j;
                uint32_t cx = 0;
                for (uint32_t x = 0; x < W; ++x) {
                    const uint64_t xw = Xj[x];
                    if ((xw >> lane) & 1ull) cand[cx + (uint32_t)__popcll(xw & lanemask_lt())] = x * 64 + lane;
                    cx += (uint32_t)__popcll(xw);
                }
                sync_tree<true>();
                if (cx == 0) {
                    if (k >= h0 && k - h0 <= p) {
                        const uint64_t idx = (uint64_t)p * a.bstride + (k - h0);
                        if (a.binom_ovf[idx]) w.run_ovf = 1;
                        add_checked(w.run, a.binom[idx], w.run_ovf);
                    }
                    break;
                }
                if (h0 + p + cx < k) break;
                if (cx <= 64 || j + 1 == a.chain) {  // the rest of the chain is one entry
                    endj = j;
                    break;
                }
                uint32_t best_s = 0, best_j = INF32;
                uint64_t sum_s = 0;
                for (uint32_t c = lane; c < cx; c += 64) {
                    const uint32_t v = cand[c];
                    const uint64_t *rv = rows + (uint64_t)v * W;
                    uint32_t sc = 0;
                    for (uint32_t x = 0; x < W; ++x) sc += (uint32_t)__popcll(Xj[x] & rv[x]);
                    sum_s += sc;
                    if (best_j == INF32 || sc > best_s) {
                        best_s = sc;
                        best_j = v;
                    }
                }
                uint64_t key = best_j != INF32 ? (((uint64_t)best_s << 32) | (INF32 - best_j)) : 0ull;
                key = wave_max64(key);
                uint32_t dummy = 0;
                const uint64_t tot_s = wave_sum64_checked(sum_s, dummy);
                const uint32_t max_s = (uint32_t)(key >> 32), pv = INF32 - (uint32_t)(key & 0xFFFFFFFFull);
                if (j == 0 && h0 + 2 == k) {  // k-cliques of T with the root's h0 held: |E(T)|
                    add_checked(w.run, tot_s / 2, w.run_ovf);
                    break;
                }
                if (h0 + p + max_s + 1 < k) break;
                const uint64_t *rp = rows + (uint64_t)pv * W;
                uint64_t *Xn = chX + (uint64_t)(j + 1) * 2 * a.wmax;
                for (uint32_t x = 0; x < W; ++x) {
                    uint64_t bw = Xj[x] & ~rp[x];
                    if (x == (pv >> 6)) bw &= ~(1ull << (pv & 63));
                    if (lane == 0) {
                        Bj[x] = bw;
                        Xn[x] = Xj[x] & ~bw & rp[x];
                    }
                    if ((bw >> lane) & 1ull) offs[ne + (uint32_t)__popcll(bw & lanemask_lt())] = (j << 16) | (x * 64 + lane);
                    ne += (uint32_t)__popcll(bw);
                }
                sync_tree<true>();
            }
            if (lane == 0) {
                if (endj != INF32) offs[ne++] = (endj << 16) | 0xFFFFu;
                ctl[2] = 0;
                ctl[3] = ne;
            }
        }
        __syncthreads();
        const uint32_t ne = ctl[3];
        for (;;) {
            uint32_t i = 0;
            if (lane == 0) i = atomicAdd(&ctl[2], 1u);
            i = uni32(i);
            if (i >= ne) break;
            const uint32_t e = offs[i], j = e >> 16, b = e & 0xFFFFu;
            const uint64_t *Xj = chX + (uint64_t)j * 2 * a.wmax, *Bj = chB + (uint64_t)j * 2 * a.wmax;
            const bool end = (b == 0xFFFFu);
            const uint64_t *rb = rows + (uint64_t)(end ? 0 : b) * W;
            for (uint32_t x = lane; x < W; x += 64) {
                if (end) {
                    X[x] = Xj[x];
                } else {
                    // the non-pivot branches of node j taken before b
                    const uint64_t taken = x < (b >> 6) ? Bj[x] : x == (b >> 6) ? (Bj[x] & ((1ull << (b & 63)) - 1)) : 0ull;
                    X[x] = Xj[x] & ~taken & rb[x];
                }
            }
            const uint32_t h = end ? h0 : h0 + 1;
            add_checked(w.run,
                        pivot_count<false, true>(a, g.psize, k, rows, stk, X, scal, cand, R, W, h, w.nodes, w.run_ovf,
                                                 nullptr, j, false, crows2, crows4),
                        w.run_ovf);
        }
        __syncthreads();
        if (wv == 0) {
            const uint64_t dt = wall_clock64() - t0;
            w.tmax = dt > w.tmax ? dt : w.tmax;
            w.tsum += dt;
        }
    }
    finish_wave(a, w);
}

// ---- task sizes and classes over the concatenated task space of a launch ----
// Query q owns global task ids [qoff[q], qoff[q+1]): its vertices (node
// granularity) or its oriented edges (edge granularity).
__device__ inline uint32_t query_of(const uint64_t *qoff, uint32_t nq, uint64_t g)
{
    uint32_t lo = 0, hi = nq - 1;  // last q with qoff[q] <= g
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (qoff[mid] <= g) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// size of a task: d+(u) for a node, |T| = |N+(u) & N+(v)| for an edge (merge)
__global__ __launch_bounds__(256) void k_task_size(const QDesc *qd, const uint64_t *qoff, uint32_t nq, uint64_t total,
                                                    uint32_t *size)
{
    for (uint64_t gi = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; gi < total;
         gi += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t q = query_of(qoff, nq, gi);
        const QDesc d = qd[q];
        const uint32_t t = (uint32_t)(gi - qoff[q]);
        if (!d.gran) {
            size[gi] = d.rowptr[t + 1] - d.rowptr[t];
            continue;
        }
        const uint32_t u = d.src[t], v = d.col[t];
        uint32_t i = d.rowptr[u], ie = d.rowptr[u + 1], j = d.rowptr[v], je = d.rowptr[v + 1];
        uint32_t c = 0;
        while (i < ie && j < je) {
            const uint32_t x = d.col[i], y = d.col[j];
            c += (x == y);
            i += (x <= y);
            j += (y <= x);
        }
        size[gi] = c;
    }
}

// size class of a task; for pivoting, tasks above top2 (<= 256) go to the
// block-wide kernel instead of the four-word register kernel
__device__ inline int class_of(uint32_t s, uint32_t top2)
{
    return s <= 64 ? 0 : s <= 128 ? 1 : s <= top2 ? 2 : s <= 512 ? 3 : s <= 1024 ? 4 : 5;
}

// Bin tasks by (algorithm, size class, size slice); a task smaller than the
// clique part still to find (k-1 for a node, k-2 for an edge) is dropped.
// Each class from 65 to 1024 vertices is cut into NSUB slices of its size
// range, largest first, and a class's slices are contiguous in the task
// array, so a class launch drains its largest tasks first and the long ones
// do not start last.  Pass 0 counts per slice and takes the max size of the
// unbounded class and of the launch (bin_max[15]), pass 1 scatters
// (query << 32 | task).  A wave handles the distinct slices among its lanes
// one at a time (ballot per slice, one atomic per slice).
constexpr int NSUB = 16;

__device__ inline int slice_of(uint32_t s, int c, uint32_t top2)
{
    if (c == 0 || c == NCLASS - 1) return 0;  // (the many small tasks: one slice keeps classification cheap)
    const uint32_t lo = c == 3 ? top2 : CLASS_ROWS[c - 1], hi = c == 2 ? top2 : CLASS_ROWS[c];
    const uint32_t sl = (hi - s) * NSUB / (hi - lo);
    return sl < NSUB ? (int)sl : NSUB - 1;
}

__global__ __launch_bounds__(256) void k_classify(const QDesc *qd, const uint8_t *qalgo, const uint64_t *qoff,
                                                   uint32_t nq, uint64_t total, const uint32_t *size, uint32_t *sub_cnt,
                                                   uint32_t *bin_max, uint32_t *sub_pos, const uint32_t *sub_off,
                                                   uint64_t *out, int scatter, uint32_t pivot_top2)
{
    const int lane = threadIdx.x & 63;
    for (uint64_t base = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) & ~63ull; base < total;
         base += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t gi = base + lane;
        int fb = -1;
        uint32_t s = 0, q = 0;
        if (gi < total) {
            q = query_of(qoff, nq, gi);
            const QDesc d = qd[q];
            const uint32_t need = d.gran ? d.k - 2 : d.k - 1;
            s = size[gi];
            if (s >= need && (d.sample <= 1 || (uint32_t)(gi - qoff[q]) % d.sample == d.offset)) {
                const uint32_t top2 = qalgo[q] ? pivot_top2 : CLASS_ROWS[2];
                const int c = class_of(s, top2);
                fb = ((qalgo[q] ? NCLASS : 0) + c) * NSUB + slice_of(s, c, top2);
            }
        }
        if (!scatter) {  // the largest task size of the launch (binomial table rows)
            uint32_t mx = fb >= 0 ? s : 0;
            for (int dd = 32; dd >= 1; dd >>= 1) mx = max(mx, (uint32_t)__shfl_xor(mx, dd, 64));
            if (lane == 0 && mx) atomicMax(&bin_max[15], mx);
        }
        for (uint64_t pend = __ballot(fb >= 0); pend;) {
            const int leader = __ffsll((unsigned long long)pend) - 1;
            const int cur = __builtin_amdgcn_readlane(fb, leader);
            const uint64_t bal = __ballot(fb == cur);
            pend &= ~bal;
            if (!scatter) {
                if (lane == leader) atomicAdd(&sub_cnt[cur], (uint32_t)__popcll(bal));
                if ((cur / NSUB) % NCLASS == NCLASS - 1) {  // the unbounded class sizes its rows by its max
                    uint32_t mx = (fb == cur) ? s : 0;
                    for (int dd = 32; dd >= 1; dd >>= 1) mx = max(mx, (uint32_t)__shfl_xor(mx, dd, 64));
                    if (lane == leader) atomicMax(&bin_max[cur / NSUB], mx);
                }
            } else {
                uint32_t b0 = 0;
                if (lane == leader) b0 = atomicAdd(&sub_pos[cur], (uint32_t)__popcll(bal));
                b0 = __shfl(b0, leader, 64);
                if (fb == cur)
                    out[sub_off[cur] + b0 + (uint32_t)__popcll(bal & lanemask_lt())] =
                        ((uint64_t)q << 32) | (uint32_t)(gi - qoff[q]);
            }
        }
    }
}

// ---- host -------------------------------------------------------------------
static const void *generic_kernel(bool pivot, bool global, bool local)
{
    static const void *const k[8] = {
        reinterpret_cast<const void *>(&k_count<false, false, false>),
        reinterpret_cast<const void *>(&k_count<false, true, false>),
        reinterpret_cast<const void *>(&k_count<true, false, false>),
        reinterpret_cast<const void *>(&k_count<true, true, false>),
        reinterpret_cast<const void *>(&k_count<false, false, true>),
        reinterpret_cast<const void *>(&k_count<false, true, true>),
        reinterpret_cast<const void *>(&k_count<true, false, true>),
        reinterpret_cast<const void *>(&k_count<true, true, true>)};
    return k[(local ? 4 : 0) + (pivot ? 2 : 0) + (global ? 1 : 0)];
}

static const void *register_kernel(int c)
{
    static const void *const reg[3] = {reinterpret_cast<const void *>(&k_pivot_reg<1>),
                                       reinterpret_cast<const void *>(&k_pivot_reg<2>),
                                       reinterpret_cast<const void *>(&k_pivot_reg<4>)};
    return reg[c];
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

__global__ __launch_bounds__(256) void k_fill_u32(uint32_t *p, uint64_t n, uint32_t v)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// The HIP events of one run_queries call, owned here so that every return
// path destroys them.  With a fixed scratch (no allocation of any kind) there
// are none: records and elapsed times become no-ops and the launches stay on
// one stream (the side-by-side launches need events to join).
struct RunEvents {
    static constexpr int NB = 2 * NCLASS;
    hipEvent_t start = nullptr, cls = nullptr, go = nullptr, fan = nullptr, end = nullptr;
    hipEvent_t b0[NB] = {}, b1[NB] = {};
    bool on = false;
    int create()
    {
        on = true;
        hipEvent_t *all[5] = {&start, &cls, &go, &fan, &end};
        for (hipEvent_t *e : all) KC_HIP(hipEventCreate(e));
        for (int b = 0; b < NB; ++b) {
            KC_HIP(hipEventCreate(&b0[b]));
            KC_HIP(hipEventCreate(&b1[b]));
        }
        return KC_OK;
    }
    void record(hipEvent_t e, hipStream_t st) const
    {
        if (on) hipEventRecord(e, st);
    }
    float ms(hipEvent_t a, hipEvent_t b) const
    {
        float t = 0;
        if (on) hipEventElapsedTime(&t, a, b);
        return t;
    }
    ~RunEvents()
    {
        hipEvent_t all[5] = {start, cls, go, fan, end};
        for (hipEvent_t e : all)
            if (e) hipEventDestroy(e);
        for (int b = 0; b < NB; ++b) {
            if (b0[b]) hipEventDestroy(b0[b]);
            if (b1[b]) hipEventDestroy(b1[b]);
        }
    }
};

int run_queries(Scratch &scr, hipStream_t s, const std::vector<QueryRun> &qs, std::vector<kc_result> &res,
                LaunchStats *ls, uint64_t *local, uint32_t grid_share)
{
    DeviceCtx *ctx;
    int rc = device_ctx(&ctx);
    if (rc) return rc;
    const uint32_t nq = (uint32_t)qs.size();
    res.assign(nq, kc_result{});
    if (nq == 0) return KC_OK;
    if (local && nq != 1) return fail(KC_ERR_ARG, "per-vertex counts take one query per launch");
    constexpr int NB = 2 * NCLASS;  // (algo, size class) bins

    // --- query descriptors and the concatenated task space ---
    std::vector<QDesc> hq(nq);
    std::vector<uint8_t> halgo(nq);
    std::vector<uint64_t> hoff(nq + 1, 0);
    uint32_t kmax = 0, P = 1;
    for (uint32_t i = 0; i < nq; ++i) {
        const kc_oriented *o = qs[i].o;
        const kc_query &q = qs[i].q;
        hq[i] = QDesc{o->rowptr.as<uint32_t>(), o->col.as<uint32_t>(), o->src.as<uint32_t>(), q.k, q.gran,
                      std::max<uint32_t>(1, qs[i].sample), qs[i].sample > 1 ? qs[i].offset : 0, q.part_size};
        halgo[i] = (uint8_t)(q.algo == KC_ALGO_PIVOT);
        hoff[i + 1] = hoff[i] + (q.gran == KC_GRAN_EDGE ? o->m : o->n);
        kmax = std::max(kmax, q.k);
        // every query keeps its own partition size; the orientation DFS stacks
        // are laid out for the smallest (the most partitions per wave)
        P = i == 0 ? q.part_size : std::min(P, q.part_size);
    }
    const uint64_t total = hoff[nq];
    if (total >= (1ull << 32) * 64) return fail(KC_ERR_LIMIT, "too many tasks in one launch");
    const size_t qd_bytes = sizeof(QDesc) * nq, off_bytes = sizeof(uint64_t) * (nq + 1);
    constexpr int NS = NB * NSUB;  // size slices
    // sub_cnt[NS] bin_max[16] sub_pos[NS] sub_off[NS] qhead[16]
    const size_t meta_u32 = 3 * NS + 32;
    const size_t acc_bytes = sizeof(uint64_t) * NB * ACC;
    const size_t q_bytes = sizeof(uint64_t) * nq + sizeof(uint32_t) * nq;
    if ((rc = scr.qdesc.ensure(qd_bytes + off_bytes + nq + 16)) ||
        (rc = scr.meta.ensure(meta_u32 * 4 + acc_bytes + q_bytes + 16)) ||
        (rc = scr.size.ensure(sizeof(uint32_t) * (total + 1))))
        return rc;
    QDesc *dq = scr.qdesc.as<QDesc>();
    uint64_t *doff = reinterpret_cast<uint64_t *>(scr.qdesc.as<uint8_t>() + qd_bytes);
    uint8_t *dalgo = scr.qdesc.as<uint8_t>() + qd_bytes + off_bytes;
    uint32_t *m32 = scr.meta.as<uint32_t>();
    uint32_t *sub_cnt = m32, *bin_max = m32 + NS, *sub_pos = bin_max + 16, *sub_off = sub_pos + NS, *qhead = sub_off + NS;
    uint64_t *dacc = reinterpret_cast<uint64_t *>(m32 + meta_u32);
    uint64_t *qcount = dacc + NB * ACC;
    uint32_t *qovf = reinterpret_cast<uint32_t *>(qcount + nq);
    KC_HIP(hipMemcpyAsync(dq, hq.data(), qd_bytes, hipMemcpyHostToDevice, s));
    KC_HIP(hipMemcpyAsync(doff, hoff.data(), off_bytes, hipMemcpyHostToDevice, s));
    KC_HIP(hipMemcpyAsync(dalgo, halgo.data(), nq, hipMemcpyHostToDevice, s));
    {  // counters, queue heads and per-query sums start at zero (a fill kernel, no memset node)
        const uint64_t nz = (meta_u32 * 4 + acc_bytes + q_bytes) / 4;
        hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)std::min<uint64_t>(1024, (nz + 255) / 256)), dim3(256), 0, s,
                           m32, nz, 0u);
    }

    // largest pivoting task the four-word register kernel takes (the rest of
    // the 129..256 class goes to the block-wide kernel); KC_PIVOT_TOP2
    // overrides it (129..256)
    uint32_t pivot_top2 = PIVOT_TOP2;
    if (const char *e = std::getenv("KC_PIVOT_TOP2")) {
        const long v = std::strtol(e, nullptr, 10);
        if (v >= 129 && v <= 256) pivot_top2 = (uint32_t)v;
    }
    RunEvents evs;
    if (!scr.fixed && (rc = evs.create())) return rc;
    evs.record(evs.start, s);
    const int gs = (int)std::max<uint64_t>(1, std::min<uint64_t>((total + 255) / 256, (uint64_t)ctx->cu * 16));
    // task sizes of a single query on the graph they were last computed for are reused
    const bool sizes_cached = nq == 1 && scr.size_uid == qs[0].o->uid && scr.size_gran == qs[0].q.gran &&
                              scr.size_total == total;
    scr.size_uid = 0;
    if (total > 0) {
        if (!sizes_cached)
            hipLaunchKernelGGL(k_task_size, dim3(gs), dim3(256), 0, s, dq, doff, nq, total, scr.size.as<uint32_t>());
        hipLaunchKernelGGL(k_classify, dim3(gs), dim3(256), 0, s, dq, dalgo, doff, nq, total, scr.size.as<uint32_t>(),
                           sub_cnt, bin_max, sub_pos, sub_off, scr.tasks.as<uint64_t>(), 0, pivot_top2);
    }
    KC_HIP(hipGetLastError());
    scr.size_uid = nq == 1 ? qs[0].o->uid : 0;
    scr.size_gran = qs[0].q.gran;
    scr.size_total = total;
    uint32_t h[NS + 16];
    KC_HIP(hipMemcpyAsync(h, m32, sizeof h, hipMemcpyDeviceToHost, s));
    KC_HIP(hipStreamSynchronize(s));
    uint32_t bcnt[NB], bmax[NB], boff[NB], soff[NS];
    uint64_t tot = 0;
    uint32_t maxsize = 0;
    for (int b = 0; b < NB; ++b) {
        boff[b] = (uint32_t)tot;
        bcnt[b] = 0;
        for (int sl = 0; sl < NSUB; ++sl) {
            soff[b * NSUB + sl] = (uint32_t)tot;
            bcnt[b] += h[b * NSUB + sl];
            tot += h[b * NSUB + sl];
        }
        bmax[b] = h[NS + b];  // (kept for the unbounded classes only)
    }
    maxsize = h[NS + 15];
    if (tot >= 0xFFFFFFFFull) return fail(KC_ERR_LIMIT, "more than 2^32 tasks in one launch");
    if ((rc = scr.tasks.ensure(sizeof(uint64_t) * ((size_t)tot + 1)))) return rc;
    KC_HIP(hipMemcpyAsync(sub_off, soff, sizeof soff, hipMemcpyHostToDevice, s));
    if (tot > 0)
        hipLaunchKernelGGL(k_classify, dim3(gs), dim3(256), 0, s, dq, dalgo, doff, nq, total, scr.size.as<uint32_t>(),
                           sub_cnt, bin_max, sub_pos, sub_off, scr.tasks.as<uint64_t>(), 1, pivot_top2);
    KC_HIP(hipGetLastError());
    evs.record(evs.cls, s);

    // --- binomial table for the pivot leaves ---
    const uint32_t bstride = kmax + 1;
    const size_t nbin = (size_t)(maxsize + 2) * bstride;
    {
        std::vector<uint64_t> bv;
        std::vector<uint8_t> bo;
        binom_table(maxsize + 1, kmax, bv, bo);
        if ((rc = scr.binom.ensure(nbin * 9 + 16))) return rc;
        KC_HIP(hipMemcpyAsync(scr.binom.p, bv.data(), nbin * 8, hipMemcpyHostToDevice, s));
        KC_HIP(hipMemcpyAsync(scr.binom.as<uint8_t>() + nbin * 8, bo.data(), nbin, hipMemcpyHostToDevice, s));
        KC_HIP(hipStreamSynchronize(s));  // the host vectors go out of scope
    }

    // --- one persistent launch per non-empty bin, largest tasks first ---
    // The bins of tasks of at most 256 vertices can run side by side on the
    // device's side streams (each bin has its own queue head and
    // accumulators), so the waves a bin frees while its last long tasks
    // finish take the next bin's tasks.  Larger classes (the wide kernel,
    // launches that use the shared HBM slice gscr) run first and in order on
    // s, and the side streams start after them: sharing the device with the
    // small classes delays the wide classes' long chain walks.  With
    // grid_share > 1 (a server worker) every launch stays on s.
    // an error after the first launch waits for every launch before it
    // returns, so that no kernel still runs on the caller's scratch
    auto drain = [&]() {
        hipStreamSynchronize(s);
        for (hipStream_t st : ctx->side) hipStreamSynchronize(st);
    };
#define KC_HIP_L(call)                                                                              \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess) {                                                                    \
            drain();                                                                               \
            return kc::fail(KC_ERR_DEVICE, "%s failed: %s", #call, hipGetErrorString(e_));         \
        }                                                                                          \
    } while (0)
    evs.record(evs.go, s);
    // Side by side only when the register classes hold few tasks per CU: then
    // a class's last long tasks are a large part of its launch (Chung-Lu
    // n=1.7M m=11.1M, edge tasks at k=8: 8,600 per CU, one 125 ms task in a
    // 132 ms launch; the count 572 -> 526 ms).  With many tasks per CU the tail
    // is negligible and overlapping launches only contend (R-MAT scale 22,
    // edge tasks at k=5: 166,000 per CU, 3.69 -> 3.91 s).
    constexpr uint64_t FAN_MAX_TASKS_PER_CU = 1u << 15;
    uint64_t reg_tasks = 0;
    for (int b = 0; b < NB; ++b)
        if (b % NCLASS <= 2) reg_tasks += bcnt[b];
    // KC_SIDE_BY_SIDE=1 / =0 forces the choice (tests run both ways)
    const char *fan_env = std::getenv("KC_SIDE_BY_SIDE");
    const bool fan_rule = reg_tasks < FAN_MAX_TASKS_PER_CU * (uint64_t)ctx->cu;
    const bool fan = evs.on && grid_share == 1 && (fan_env && *fan_env ? *fan_env == '1' : fan_rule);
    bool side_waits[DeviceCtx::NSIDE] = {false}, fan_recorded = false;
    hipStream_t bin_stream[NB];
    int nside = 0;
    uint32_t variant[NB] = {0};
    bool launched[NB] = {false};
    const size_t lds_cap = ctx->lds_per_block;
    int order[NB], no = 0;
    for (int c = NCLASS - 1; c >= 0; --c) {
        order[no++] = NCLASS + c;  // pivot bins
        order[no++] = c;           // orientation bins
    }
    for (int oi = 0; oi < NB; ++oi) {
        const int b = order[oi];
        if (bcnt[b] == 0) continue;
        const bool pivot = b >= NCLASS;
        const int c = b % NCLASS;
        CountArgs a{};
        a.qd = dq;
        a.tasks = scr.tasks.as<uint64_t>() + boff[b];
        a.ntasks = bcnt[b];
        a.binom = scr.binom.as<uint64_t>();
        a.binom_ovf = scr.binom.as<uint8_t>() + nbin * 8;
        a.bstride = bstride;
        a.qhead = qhead + b;
        a.acc = dacc + ACC * b;
        a.qcount = qcount;
        a.qovf = qovf;
        a.local = local;
        const bool lc = local != nullptr;
        bool global = (c == NCLASS - 1);
        const uint32_t rows_cap = global ? bmax[b] : CLASS_ROWS[c];
        a.rmax = rows_cap;
        a.wmax = (rows_cap + 63) / 64;
        a.levels = pivot ? kmax + 3 : std::max<uint32_t>(1, kmax - 2);
        uint32_t np = 64 / P;
        Layout L = make_layout(a.rmax, a.wmax, a.levels, np, pivot, lc);
        if (!global && !pivot) {
            while (np > 1 && L.total * 8 > lds_cap) {  // fewer partitions when the stacks do not fit
                np >>= 1;
                L = make_layout(a.rmax, a.wmax, a.levels, np, pivot, lc);
            }
        }
        if (!global && L.total * 8 > lds_cap) global = true;  // same layout in HBM
        a.nparts = np;
        const void *kern = generic_kernel(pivot, global, lc);
        variant[b] = global ? 3u : 1u;
        if (pivot && !global && c <= 2 && !lc) {
            kern = register_kernel(c);
            L = make_layout(a.rmax, a.wmax, 0, 0, true);
            variant[b] = 2u;
        }
        // large tasks of pivoting: blocks of several waves sharing one bitmap
        uint32_t nw = 1;
        WideLayout WL{};
        if (pivot && !lc && c >= 3) {
            // chain levels walked by wave 0 before the entries are shared: the
            // whole chain for the HBM classes (a chain left whole is one entry,
            // which one wave would walk alone), 32 levels when LDS holds them
            if (a.rmax >= 0xFFFFu) {
                drain();
                return fail(KC_ERR_LIMIT, "a task of %u vertices exceeds the wide kernel's 65534", a.rmax);
            }
            a.chain = c == 3 ? 32u : std::min<uint32_t>(1024u, std::max<uint32_t>(64u, a.rmax));
            bool fits = false;
            if (c < NCLASS - 1)
                for (uint32_t t = 8; t >= 2 && !fits; t >>= 1) {
                    WL = make_wide_layout(a.rmax, a.wmax, a.levels, t, a.chain);
                    if (WL.total * 8 <= lds_cap) {
                        nw = t;
                        fits = true;
                    }
                }
            global = !fits;
            if (!fits) {
                nw = 8;
                WL = make_wide_layout(a.rmax, a.wmax, a.levels, nw, a.chain);
            }
            kern = global ? reinterpret_cast<const void *>(&k_pivot_wide<true>)
                          : reinterpret_cast<const void *>(&k_pivot_wide<false>);
            L.total = WL.total;
            variant[b] = global ? 5u : 4u;
        }
        const size_t lds = global ? 0 : L.total * 8;
        if (!global) KC_HIP_L(hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        int per_cu = 0;
        KC_HIP_L(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, (int)(64 * nw), lds));
        if (per_cu < 1) per_cu = 1;
        uint64_t blocks = std::max<uint64_t>(1, (uint64_t)per_cu * ctx->cu / std::max<uint32_t>(1, grid_share));
        a.chunk = (c == 0 && bcnt[b] > blocks * 64) ? 4 : 1;  // (the wide kernel takes one task per grab)
        blocks = std::min<uint64_t>(blocks, ((uint64_t)bcnt[b] + a.chunk - 1) / a.chunk);
        if (global) {
            a.gscr_words = L.total;
            // a fixed scratch holds as many slices as fit in what it has left
            const uint64_t cap = (scr.fixed ? scr.gscr.bytes : (2ull << 30)) / (L.total * 8 + 1);
            if (scr.fixed && cap == 0) {
                drain();
                return fail(KC_ERR_LIMIT, "a task of %u vertices needs %llu bytes of HBM workspace, more than the "
                            "%zu left in the fixed scratch", a.rmax, (unsigned long long)(L.total * 8), scr.gscr.bytes);
            }
            blocks = std::max<uint64_t>(1, std::min<uint64_t>(blocks, cap));
            if ((rc = scr.gscr.ensure(blocks * L.total * 8))) {
                drain();
                return rc;
            }
            a.gscr = scr.gscr.as<uint64_t>();
        }
        hipStream_t st = s;
        if (fan && !global && c <= 2) {
            if (!fan_recorded) {
                KC_HIP_L(hipEventRecord(evs.fan, s));  // after every launch so far on s
                fan_recorded = true;
            }
            const int si = nside++ % DeviceCtx::NSIDE;
            st = ctx->side[si];
            if (!side_waits[si]) {
                KC_HIP_L(hipStreamWaitEvent(st, evs.fan, 0));
                side_waits[si] = true;
            }
        }
        bin_stream[b] = st;
        void *args[] = {&a};
        evs.record(evs.b0[b], st);
        KC_HIP_L(hipLaunchKernel(kern, dim3((unsigned)blocks), dim3(64 * nw), args, lds, st));
        evs.record(evs.b1[b], st);
        launched[b] = true;
    }
    for (int b = 0; b < NB; ++b)
        if (launched[b] && bin_stream[b] != s) KC_HIP_L(hipStreamWaitEvent(s, evs.b1[b], 0));
    evs.record(evs.end, s);
    std::vector<uint64_t> hq_count(nq);
    std::vector<uint32_t> hq_ovf(nq);
    uint64_t hacc[NB * ACC];
    KC_HIP_L(hipMemcpyAsync(hacc, dacc, sizeof hacc, hipMemcpyDeviceToHost, s));
    KC_HIP_L(hipMemcpyAsync(hq_count.data(), qcount, sizeof(uint64_t) * nq, hipMemcpyDeviceToHost, s));
    KC_HIP_L(hipMemcpyAsync(hq_ovf.data(), qovf, sizeof(uint32_t) * nq, hipMemcpyDeviceToHost, s));
    const hipError_t se = hipStreamSynchronize(s);
    if (se != hipSuccess) return fail(KC_ERR_DEVICE, "counting kernels failed: %s", hipGetErrorString(se));
#undef KC_HIP_L
    float bin_ms[NB] = {0};
    const float prep_ms = evs.ms(evs.start, evs.cls), total_ms = evs.ms(evs.go, evs.end);
    for (int b = 0; b < NB; ++b)
        if (launched[b]) bin_ms[b] = evs.ms(evs.b0[b], evs.b1[b]);
    int wall_khz = 100000;
    hipDeviceGetAttribute(&wall_khz, hipDeviceAttributeWallClockRate, ctx->dev);
    if (wall_khz <= 0) wall_khz = 100000;

    for (uint32_t i = 0; i < nq; ++i) {
        kc_result &r = res[i];
        r.count = hq_count[i];
        r.overflow = hq_ovf[i] ? 1 : 0;
        r.algo = qs[i].q.algo;
        r.gran = qs[i].q.gran;
        r.part_size = qs[i].q.part_size;
    }
    if (ls) {
        *ls = LaunchStats{};
        ls->count_ms = total_ms;
        ls->prep_ms = prep_ms;
        int dom = -1;
        for (int b = 0; b < NB; ++b) {
            ls->tasks += hacc[ACC * b + 2];
            ls->alg_bytes += hacc[ACC * b + 1];
            ls->tree_nodes += hacc[ACC * b + 3];
            // the dominant launch: the one whose tasks hold the waves longest
            // (summed task durations; the side-by-side launches' spans overlap)
            if (launched[b] && (dom < 0 || hacc[ACC * b + 5] > hacc[ACC * dom + 5])) dom = b;
            const int c = b % NCLASS;
            if (!launched[b]) continue;
            ls->class_ms[c] += bin_ms[b];
            ls->class_bytes[c] += hacc[ACC * b + 1];
            ls->class_tasks[c] += hacc[ACC * b + 2];
            ls->class_kernel[c] = variant[b];
            ls->class_max_task_ms[c] = std::max(ls->class_max_task_ms[c], (double)hacc[ACC * b + 4] / wall_khz);
            ls->class_task_ms_sum[c] += (double)hacc[ACC * b + 5] / wall_khz;
        }
        if (dom >= 0) {
            ls->dominant_ms = bin_ms[dom];
            ls->dominant_bytes = hacc[ACC * dom + 1];
            ls->dominant_class = (uint32_t)(dom % NCLASS);
        }
    }
    return KC_OK;
}


// ---- OpenEngine (host) --------------------------------------------------------
OpenEngine::~OpenEngine()
{
    if (dev >= 0) hipSetDevice(dev);
    // every launch is stopped before any wait: streams sharing a hardware
    // queue would otherwise wait on a launch not stopped yet
    for (int j = 0; j < NQ; ++j)
        if (running(j) && !qs[j].stopped) stop(j);
    for (int j = 0; j < NQ; ++j) {
        if (qs[j].run) hipStreamSynchronize(qs[j].run);
        if (qs[j].done) hipEventDestroy(qs[j].done);
        if (qs[j].run) hipStreamDestroy(qs[j].run);
    }
    if (adm) {
        hipStreamSynchronize(adm);
        hipStreamDestroy(adm);
    }
    if (pinned) hipHostFree(pinned);
}

int OpenEngine::init(int device)
{
    dev = device;
    KC_HIP(hipSetDevice(dev));
    DeviceCtx *ctx;
    int rc = device_ctx(&ctx);
    if (rc) return rc;
    KC_HIP(hipStreamCreate(&adm));
    for (Queue &q : qs) {
        KC_HIP(hipStreamCreate(&q.run));
        KC_HIP(hipEventCreateWithFlags(&q.done, hipEventDisableTiming));
    }
    if ((rc = qcount.ensure(8 * SLOTS)) || (rc = qovf.ensure(4 * SLOTS)) || (rc = qdone.ensure(8 * SLOTS)) ||
        (rc = ctl.ensure(8 * CTL_WORDS * NQ)) || (rc = acc.ensure(8 * ACC)))
        return rc;
    const size_t qbytes = 16 + 16 * NSEG;  // total, nseg, then (base, slot|count) per segment
    const size_t bytes = NQ * qbytes + sizeof(OpenSlot) * SLOTS + sizeof(OpenMail) * SLOTS;
    KC_HIP(hipHostMalloc((void **)&pinned, bytes, hipHostMallocCoherent | hipHostMallocMapped));
    std::memset(pinned, 0, bytes);
    KC_HIP(hipHostGetDevicePointer((void **)&pinned_dev, pinned, 0));
    for (int j = 0; j < NQ; ++j) {
        uint64_t *h = reinterpret_cast<uint64_t *>(pinned + j * qbytes);
        qs[j].htotal = h;
        qs[j].hnseg = h + 1;
        qs[j].hseg = h + 2;
    }
    hslot = reinterpret_cast<OpenSlot *>(pinned + NQ * qbytes);
    mail = reinterpret_cast<OpenMail *>(hslot + SLOTS);
    // C(n, r) for n <= OPEN_MAX_ROWS + 1, r <= KC_MAX_K: every query a launch can take
    {
        std::vector<uint64_t> bv;
        std::vector<uint8_t> bo;
        binom_table(OPEN_MAX_ROWS + 1, KC_MAX_K, bv, bo);
        const size_t nbin = (size_t)(OPEN_MAX_ROWS + 2) * (KC_MAX_K + 1);
        if ((rc = binom.ensure(nbin * 9 + 16))) return rc;
        KC_HIP(hipMemcpyAsync(binom.p, bv.data(), nbin * 8, hipMemcpyHostToDevice, adm));
        KC_HIP(hipMemcpyAsync(binom.as<uint8_t>() + nbin * 8, bo.data(), nbin, hipMemcpyHostToDevice, adm));
        KC_HIP(hipStreamSynchronize(adm));
    }
    // per-slot sums start at zero; the flush that completes a slot zeroes them again
    hipLaunchKernelGGL(k_fill_u32, dim3(64), dim3(256), 0, adm, qcount.as<uint32_t>(), (uint64_t)2 * SLOTS, 0u);
    hipLaunchKernelGGL(k_fill_u32, dim3(64), dim3(256), 0, adm, qovf.as<uint32_t>(), (uint64_t)SLOTS, 0u);
    hipLaunchKernelGGL(k_fill_u32, dim3(64), dim3(256), 0, adm, qdone.as<uint32_t>(), (uint64_t)2 * SLOTS, 0u);
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, adm, ctl.as<uint32_t>(), (uint64_t)2 * CTL_WORDS * NQ, 0u);
    KC_HIP(hipGetLastError());
    KC_HIP(hipStreamSynchronize(adm));
    const Layout L = make_layout(OPEN_MAX_ROWS, OPEN_MAX_ROWS / 64, 0, 0, true);
    lds = L.total * 8;
    KC_HIP(hipFuncSetAttribute((const void *)k_pivot_open, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    int per_cu = 0;
    KC_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)k_pivot_open, 64, lds));
    if (per_cu < 12) return fail(KC_ERR_DEVICE, "resident launches: %d waves per CU", per_cu);
    // express: 4 waves per CU; bulk: the rest but 4, which stay free for the
    // workers' launches
    qs[0].blocks = (uint32_t)(4 * ctx->cu);
    qs[1].blocks = (uint32_t)((per_cu - 8) * ctx->cu);
    free_slots.clear();
    for (uint32_t i = SLOTS; i-- > 0;) free_slots.push_back(i);
    return KC_OK;
}

bool OpenEngine::running(int j)
{
    Queue &q = qs[j];
    if (!q.launched) return false;
    if (hipEventQuery(q.done) == hipSuccess) q.launched = false;
    return q.launched;
}

int OpenEngine::reset(int j)
{
    if (running(j) || qs[j].inflight) return fail(KC_ERR_ARG, "open queue reset while in use");
    Queue &q = qs[j];
    // the device words are cleared on the queue's own stream, ahead of its
    // next launch, and the host does not wait: more streams than hardware
    // queues share them, and a wait on a stream whose hardware queue holds
    // another resident launch would last until that launch ends -- which
    // only this thread's stop can bring about
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, q.run, ctl.as<uint32_t>() + 2 * CTL_WORDS * j,
                       (uint64_t)2 * CTL_WORDS, 0u);
    KC_HIP(hipGetLastError());
    __atomic_store_n(q.hnseg, 0ull, __ATOMIC_RELEASE);
    __atomic_store_n(q.htotal, 0ull, __ATOMIC_RELEASE);
    q.total = q.nseg = 0;
    return KC_OK;
}

static void publish_segment(OpenEngine::Queue &q, uint32_t slot, uint64_t count)
{
    q.hseg[2 * q.nseg] = q.total;
    q.hseg[2 * q.nseg + 1] = ((uint64_t)slot << 32) | count;
    q.nseg++;
    q.total += count;
}

// segments and descriptors are written before the counts that publish them;
// the segment count before the task total (a wave that sees the total finds
// its segment)
static void publish_counts(OpenEngine::Queue &q)
{
    __atomic_store_n(q.hnseg, q.nseg, __ATOMIC_RELEASE);
    __atomic_store_n(q.htotal, q.total, __ATOMIC_RELEASE);
}

void OpenEngine::admit(int j, const std::vector<OpenQueryIn> &in, std::vector<uint32_t> &slot)
{
    Queue &qq = qs[j];
    slot.assign(in.size(), ~0u);
    bool any = false;
    for (size_t i = 0; i < in.size(); ++i) {
        const OpenQueryIn &x = in[i];
        const uint64_t nt = x.q.gran == KC_GRAN_EDGE ? x.o->m : x.o->n;
        if (nt == 0) continue;
        const uint32_t sl = free_slots.back();
        free_slots.pop_back();
        slot[i] = sl;
        hslot[sl] = OpenSlot{x.o->rowptr.as<uint32_t>(), x.o->col.as<uint32_t>(), x.o->src.as<uint32_t>(), x.q.k,
                             x.q.gran, nt, x.seq, 0};
        publish_segment(qq, sl, nt);
        qq.inflight++;
        any = true;
    }
    if (any) publish_counts(qq);
}

void OpenEngine::stop(int j)
{
    Queue &q = qs[j];
    publish_segment(q, OPEN_STOP, q.blocks);
    publish_counts(q);
    q.stopped = true;
}

int OpenEngine::launch(int j)
{
    Queue &q = qs[j];
    OpenArgs oa{};
    CountArgs &a = oa.c;
    a.rmax = OPEN_MAX_ROWS;
    a.wmax = OPEN_MAX_ROWS / 64;
    a.binom = binom.as<uint64_t>();
    a.binom_ovf = binom.as<uint8_t>() + (size_t)(OPEN_MAX_ROWS + 2) * (KC_MAX_K + 1) * 8;
    a.bstride = KC_MAX_K + 1;
    a.acc = acc.as<uint64_t>();
    a.qcount = qcount.as<uint64_t>();
    a.qovf = qovf.as<uint32_t>();
    a.chunk = 1;
    uint64_t *c = ctl.as<uint64_t>() + CTL_WORDS * j;
    oa.counter = c;
    oa.known = c + 1;
    oa.last_poll = c + 2;
    oa.qdone = qdone.as<uint64_t>();
    auto dev_of = [&](const void *h) { return pinned_dev + (reinterpret_cast<const uint8_t *>(h) - pinned); };
    oa.htotal = reinterpret_cast<const uint64_t *>(dev_of(q.htotal));
    oa.hnseg = reinterpret_cast<const uint64_t *>(dev_of(q.hnseg));
    oa.hseg = reinterpret_cast<const uint64_t *>(dev_of(q.hseg));
    oa.hslot = reinterpret_cast<const OpenSlot *>(dev_of(hslot));
    oa.mail = reinterpret_cast<OpenMail *>(dev_of(mail));
    // a backstop only (the host ends every launch with a stop segment): 10 min
    oa.safety_ticks = 60000000000ull;
    void *args[] = {&oa};
    KC_HIP(hipLaunchKernel((const void *)k_pivot_open, dim3(q.blocks), dim3(64), args, lds, q.run));
    KC_HIP(hipEventRecord(q.done, q.run));
    q.launched = true;
    q.stopped = false;
    return KC_OK;
}

}  // namespace kc

using namespace kc;

namespace kc {
int check_query(const kc_query *q);
}
int kc::check_query(const kc_query *q)
{
    if (q->k < 1 || q->k > KC_MAX_K) return fail(KC_ERR_ARG, "k=%u outside 1..%d", q->k, KC_MAX_K);
    if (q->algo > 1 || q->gran > 1) return fail(KC_ERR_ARG, "bad algo/gran");
    const uint32_t P = q->part_size;
    if (P == 0 || P > 64 || (P & (P - 1))) return fail(KC_ERR_ARG, "partition size %u", P);
    return KC_OK;
}

// The autotuner (auto_tune = 1).  Measured on gfx950: one lane per candidate
// row / per DFS partition (P = 1) is fastest for every size class; pivoting
// with closed-form tails is never slower than the orientation DFS; edge tasks
// balance heavy-tailed degree distributions, node tasks win when every
// out-degree fits one 64-bit word and there are fewer vertices than edges to
// schedule.
namespace kc {
kc_query tune_query(const kc_oriented *o, const kc_query *q);
}
static kc_query tune(const kc_oriented *o, const kc_query *q) { return kc::tune_query(o, q); }

kc_query kc::tune_query(const kc_oriented *o, const kc_query *q)
{
    kc_query t = *q;
    if (q->auto_tune) {
        t.algo = KC_ALGO_PIVOT;
        t.part_size = 1;
        t.gran = (o->m > 0 && o->max_out > 64) ? KC_GRAN_EDGE : KC_GRAN_NODE;
    }
    return t;
}

void kc::fill_result(kc_result *r, const kc_result &res, const LaunchStats &ls)
{
    *r = res;
    r->tasks = ls.tasks;
    r->alg_bytes = ls.alg_bytes;
    r->tree_nodes = ls.tree_nodes;
    r->count_ms = ls.count_ms;
    r->dominant_ms = ls.dominant_ms;
    r->dominant_bytes = ls.dominant_bytes;
    r->dominant_class = ls.dominant_class;
    for (int c = 0; c < 6; ++c) {
        r->class_ms[c] = ls.class_ms[c];
        r->class_bytes[c] = ls.class_bytes[c];
        r->class_tasks[c] = ls.class_tasks[c];
        r->class_kernel[c] = ls.class_kernel[c];
        r->class_max_task_ms[c] = ls.class_max_task_ms[c];
        r->class_task_ms_sum[c] = ls.class_task_ms_sum[c];
    }
}

// ---- measured autotuner ---------------------------------------------------
// Both orientations are built and each (orientation, granularity) pair counts
// every 64th task at this k with pivoting; the pair with the least device
// time (classification + count of the sample) wins.  At k <= 5, where the
// orientation DFS can compete (its tree is at most three levels deep below
// the task), the same sample is then counted with the orientation DFS at
// partitions of 1, 8 and 64 lanes on the winning pair, and the algorithm and
// partition size with the least time make the plan.  (Summed task times,
// tail-free, ranked the plans worse: a sample of few large tasks runs at
// another occupancy than one of many small ones.)
int kc::autotune_measured(const kc_graph *g, uint32_t k, uint32_t degree_max_out, kc_plan *plan)
{
    constexpr uint32_t SAMPLE = 64;
    constexpr uint32_t ORIENT_DFS_MAX_K = 5;
    if (k > KC_MAX_K) return fail(KC_ERR_ARG, "k=%u outside 1..%d", k, KC_MAX_K);
    kc_oriented *o[2] = {nullptr, nullptr};
    int rc = kc_orient(g, KC_ORIENT_DEGREE, &o[0]);
    if (rc == KC_OK) rc = kc_orient(g, KC_ORIENT_DEGEN, &o[1]);
    double best = 0;
    int best_oi = 1;
    kc_query best_q{k, KC_ALGO_PIVOT, KC_GRAN_EDGE, 1, 1, 0};
    bool have = false;
    auto measure = [&](int oi, const kc_query &q) {
        QueryRun r;
        r.o = o[oi];
        r.q = q;
        r.sample = SAMPLE;
        std::vector<kc_result> res;
        LaunchStats ls;
        hipSetDevice(o[oi]->dev);
        rc = run_queries(o[oi]->scr, o[oi]->stream, {r}, res, &ls);
        const double t = ls.count_ms + ls.prep_ms;
        if (rc == KC_OK && (!have || t < best)) {
            have = true;
            best = t;
            best_oi = oi;
            best_q = q;
        }
    };
    for (int oi = 0; oi < 2 && rc == KC_OK; ++oi)
        for (uint32_t gran = KC_GRAN_NODE; gran <= KC_GRAN_EDGE && rc == KC_OK; ++gran)
            measure(oi, kc_query{k, KC_ALGO_PIVOT, gran, 1, 1, 0});
    if (k <= ORIENT_DFS_MAX_K && rc == KC_OK) {
        const int oi = best_oi;
        const uint32_t gran = best_q.gran;
        for (uint32_t P : {1u, 8u, 64u})
            if (rc == KC_OK) measure(oi, kc_query{k, KC_ALGO_ORIENT, gran, P, 1, 0});
    }
    for (kc_oriented *x : o)
        if (x) kc_oriented_free(x);
    if (rc != KC_OK) return rc;
    plan->degree_max_out = degree_max_out;
    plan->orient_mode = best_oi == 0 ? KC_ORIENT_DEGREE : KC_ORIENT_DEGEN;
    plan->algo = best_q.algo;
    plan->gran = best_q.gran;
    plan->part_size = best_q.part_size;
    return KC_OK;
}

// kc_count over the tasks whose id (vertex, or oriented edge in CSR order)
// is congruent to offset modulo stride; stride 1 counts every task.
static int count_tasks(const kc_oriented *o, const kc_query *q, uint32_t stride, uint32_t offset, kc_result *r)
{
    if (!o || !q || !r) return fail(KC_ERR_ARG, "kc_count: null argument");
    if (stride == 0 || offset >= stride) return fail(KC_ERR_ARG, "task sample offset %u, stride %u", offset, stride);
    int rc = check_query(q);
    if (rc) return rc;
    hipSetDevice(o->dev);
    const kc_query qq = tune(o, q);
    std::memset(r, 0, sizeof *r);
    if (qq.k <= 2) {
        // closed forms: every vertex is a 1-clique, every edge a 2-clique
        r->algo = qq.algo;
        r->gran = qq.gran;
        r->part_size = qq.part_size;
        const uint64_t ntasks = qq.gran == KC_GRAN_EDGE ? o->m : o->n;
        const uint64_t nsample = ntasks > offset ? (ntasks - offset + stride - 1) / stride : 0;
        r->tasks = nsample;
        if (qq.k == 1) {
            r->count = o->n > offset ? (o->n - offset + stride - 1) / stride : 0;
        } else if (qq.gran == KC_GRAN_EDGE || stride == 1) {
            r->count = stride == 1 ? o->m : nsample;
        } else {  // out-degrees of the sampled vertices
            std::vector<uint32_t> rp((size_t)o->n + 1);
            if ((rc = kc_oriented_export(o, rp.data(), nullptr, nullptr))) return rc;
            for (uint64_t u = offset; u < o->n; u += stride) r->count += rp[u + 1] - rp[u];
        }
        return KC_OK;
    }
    kc_oriented *om = const_cast<kc_oriented *>(o);
    QueryRun run{o, qq};
    run.sample = stride;
    run.offset = offset;
    std::vector<QueryRun> qs{run};
    std::vector<kc_result> res;
    LaunchStats ls;
    rc = run_queries(om->scr, om->stream, qs, res, &ls);
    if (rc) return rc;
    fill_result(r, res[0], ls);
    return r->overflow ? fail(KC_ERR_OVERFLOW, "k-clique count exceeds 2^64-1") : KC_OK;
}

extern "C" int kc_count(const kc_oriented *o, const kc_query *q, kc_result *r)
{
    return count_tasks(o, q, 1, 0, r);
}

extern "C" int kc_count_sampled(const kc_oriented *o, const kc_query *q, uint32_t stride, uint32_t offset,
                                kc_result *r)
{
    return count_tasks(o, q, stride, offset, r);
}

extern "C" int kc_count_local(const kc_oriented *o, const kc_query *q, uint64_t *local, kc_result *r)
{
    if (!o || !q || !r || (!local && o->n)) return fail(KC_ERR_ARG, "kc_count_local: null argument");
    int rc = check_query(q);
    if (rc) return rc;
    hipSetDevice(o->dev);
    kc_query qq = tune(o, q);
    std::memset(r, 0, sizeof *r);
    r->algo = qq.algo;
    r->gran = qq.gran;
    r->part_size = qq.part_size;
    const uint32_t n = o->n;
    if (qq.k <= 2) {
        // every vertex is in one 1-clique; in deg(v) = out + in 2-cliques
        std::vector<uint32_t> rp((size_t)n + 1), col(o->m);
        if (qq.k == 2 && (rc = kc_oriented_export(o, rp.data(), col.data(), nullptr))) return rc;
        for (uint32_t v = 0; v < n; ++v) local[v] = qq.k == 1 ? 1 : rp[v + 1] - rp[v];
        if (qq.k == 2)
            for (uint32_t x : col) local[x]++;
        r->count = qq.k == 1 ? n : o->m;
        r->tasks = qq.gran == KC_GRAN_EDGE ? o->m : o->n;
        return KC_OK;
    }
    kc_oriented *om = const_cast<kc_oriented *>(o);
    if ((rc = om->scr.local.ensure(sizeof(uint64_t) * ((size_t)n + 1)))) return rc;
    KC_HIP(hipMemsetAsync(om->scr.local.p, 0, sizeof(uint64_t) * (size_t)n, om->stream));
    std::vector<QueryRun> qs{QueryRun{o, qq}};
    std::vector<kc_result> res;
    LaunchStats ls;
    rc = run_queries(om->scr, om->stream, qs, res, &ls, om->scr.local.as<uint64_t>());
    if (rc) return rc;
    if (n) {
        KC_HIP(hipMemcpyAsync(local, om->scr.local.p, sizeof(uint64_t) * (size_t)n, hipMemcpyDeviceToHost, om->stream));
        KC_HIP(hipStreamSynchronize(om->stream));
    }
    fill_result(r, res[0], ls);
    return r->overflow ? fail(KC_ERR_OVERFLOW, "k-clique count exceeds 2^64-1") : KC_OK;
}

extern "C" int kc_batch_count(kc_oriented *const *gs, const kc_query *qs, size_t nq, kc_result *results,
                              kc_batch_stats *stats)
{
    if ((!gs || !qs || !results) && nq) return fail(KC_ERR_ARG, "kc_batch_count: null argument");
    std::vector<QueryRun> runs;
    std::vector<size_t> idx;
    runs.reserve(nq);
    int dev = -1;
    for (size_t i = 0; i < nq; ++i) {
        if (!gs[i]) return fail(KC_ERR_ARG, "kc_batch_count: null graph %zu", i);
        int rc = check_query(&qs[i]);
        if (rc) return rc;
        if (dev < 0) dev = gs[i]->dev;
        if (gs[i]->dev != dev) return fail(KC_ERR_ARG, "kc_batch_count: graphs on different devices");
        const kc_query t = tune(gs[i], &qs[i]);
        std::memset(&results[i], 0, sizeof(kc_result));
        results[i].algo = t.algo;
        results[i].gran = t.gran;
        results[i].part_size = t.part_size;
        if (t.k <= 2) {
            results[i].count = t.k == 1 ? gs[i]->n : gs[i]->m;
            continue;
        }
        runs.push_back(QueryRun{gs[i], t});
        idx.push_back(i);
    }
    if (stats) std::memset(stats, 0, sizeof *stats);
    if (runs.empty()) return KC_OK;
    hipSetDevice(dev);
    kc_oriented *first = runs[0].o == nullptr ? nullptr : const_cast<kc_oriented *>(runs[0].o);
    std::vector<kc_result> res;
    LaunchStats ls;
    int rc = run_queries(first->batch_scr, first->stream, runs, res, &ls);
    if (rc) return rc;
    bool ovf = false;
    for (size_t j = 0; j < runs.size(); ++j) {
        results[idx[j]].count = res[j].count;
        results[idx[j]].overflow = res[j].overflow;
        ovf |= res[j].overflow != 0;
    }
    if (stats) {
        stats->total_ms = ls.count_ms + ls.prep_ms;
        stats->count_ms = ls.count_ms;
        stats->tasks = ls.tasks;
        stats->queries = runs.size();
    }
    return ovf ? fail(KC_ERR_OVERFLOW, "a k-clique count exceeds 2^64-1") : KC_OK;
}
