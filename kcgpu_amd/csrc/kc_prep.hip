// kc_prep.hip -- device preprocessing: exclusive scan, k-core peel (degeneracy
// key), keep flags, per-row kept counts and stable per-row compaction that
// turn the symmetric CSR into the oriented CSR.
//
// Reference path replaced (src/main.cu:164-235):
//   getNodeDegree_kernel            kcore/kcore.cuh:50-59
//   findKcoreIncremental_async      kcore/kcore.cuh:442-584 (host loop, a D->H
//                                   count read per round) -> steps of five
//                                   kernels queued 32 at a time, one D->H
//                                   read per 32 steps (or, on request, one
//                                   cooperative kernel with grid barriers)
//   init (keep flags)               include/main_support.cuh:54-128
//   CUBSelect x2                    include/cub_wrappers.cuh:121-179
//   warp_detect_deleted_edges       include/main_support.cuh:133-167
//   CUBScanExclusive                include/cub_wrappers.cuh:11-64
// Here a wave (64 lanes) owns a row: it counts kept entries with a 64-bit
// ballot, and after the scan writes them in row order at ballot-prefix ranks,
// which is the stable compaction without a 2m-wide scan.
#include <hip/hip_cooperative_groups.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <thread>

#include "kc_device.h"

namespace cg = cooperative_groups;

namespace kc {

// ---------------------------------------------------------------------------
// device context / buffers
// ---------------------------------------------------------------------------
static DeviceCtx g_ctx;
static int g_state = 0;  // 0 untried, 1 ok, 2 failed
static std::string g_why;
static std::mutex g_mu;

int device_ctx(DeviceCtx **out)
{
    DeviceCtx &ctx = g_ctx;
    int &state = g_state;
    std::string &why = g_why;
    std::lock_guard<std::mutex> lk(g_mu);
    if (state == 0) {
        int cnt = 0;
        hipError_t e = hipGetDeviceCount(&cnt);
        if (e != hipSuccess || cnt == 0) {
            state = 2;
            why = std::string("no HIP device: ") + (e != hipSuccess ? hipGetErrorString(e) : "count 0");
        } else {
            int d = 0;
            hipGetDevice(&d);
            hipDeviceProp_t p;
            if (hipGetDeviceProperties(&p, d) != hipSuccess) {
                state = 2;
                why = "hipGetDeviceProperties failed";
            } else if (std::strncmp(p.gcnArchName, "gfx950", 6) != 0) {
                state = 2;
                why = std::string("device arch ") + p.gcnArchName + " is not gfx950";
            } else {
                ctx.dev = d;
                ctx.prop = p;
                ctx.cu = p.multiProcessorCount;
                ctx.lds_per_block = p.sharedMemPerBlock;
                ctx.hbm = p.totalGlobalMem;
                hipSetDevice(d);
                bool ok = ctx.stream || hipStreamCreate(&ctx.stream) == hipSuccess;
                for (hipStream_t &st : ctx.side) ok = ok && (st || hipStreamCreate(&st) == hipSuccess);
                if (!ok) {
                    state = 2;
                    why = "hipStreamCreate failed";
                } else {
                    state = 1;
                }
            }
        }
    }
    if (state != 1) return fail(KC_ERR_DEVICE, "%s", why.c_str());
    hipSetDevice(ctx.dev);
    *out = &ctx;
    return KC_OK;
}

int DevBuf::ensure(size_t want)
{
    if (!owned && p) {
        if (want <= bytes) return KC_OK;
        return fail(KC_ERR_LIMIT, "%zu bytes of scratch needed, the fixed region holds %zu", want, bytes);
    }
    if (want <= bytes && p && owned) return KC_OK;
    release();
    owned = true;
    size_t b = want ? want : 256;
    hipError_t e = hipMalloc(&p, b);
    if (e != hipSuccess) {
        p = nullptr;
        bytes = 0;
        return fail(KC_ERR_NOMEM, "hipMalloc(%zu) failed: %s", b, hipGetErrorString(e));
    }
    bytes = b;
    return KC_OK;
}

void DevBuf::release()
{
    if (p && owned) hipFree(p);
    p = nullptr;
    bytes = 0;
    owned = true;
}

// ---------------------------------------------------------------------------
// exclusive scan: tile reduce -> scan of tile sums (one block, carried) ->
// tile scan with offset.  256 threads x 8 items per tile.
// ---------------------------------------------------------------------------
constexpr int SCAN_T = 256, SCAN_I = 8, SCAN_TILE = SCAN_T * SCAN_I;

__device__ inline uint64_t wave_incl_scan(uint64_t x)
{
    int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint64_t y = __shfl_up(x, (unsigned)d, 64);
        if (lane >= d) x += y;
    }
    return x;
}

// block-wide inclusive scan over 256 threads; returns inclusive value, total via *tot
__device__ inline uint64_t block_incl_scan(uint64_t x, uint64_t *sh, uint64_t *tot)
{
    int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    uint64_t inc = wave_incl_scan(x);
    if (lane == 63) sh[wid] = inc;
    __syncthreads();
    uint64_t off = 0, t = 0;
    for (int w = 0; w < SCAN_T / 64; ++w) {
        if (w < wid) off += sh[w];
        t += sh[w];
    }
    __syncthreads();
    *tot = t;
    return inc + off;
}

__global__ __launch_bounds__(SCAN_T) void k_scan_reduce(const uint32_t *in, uint32_t n, uint64_t *part)
{
    __shared__ uint64_t sh[SCAN_T / 64];
    uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE;
    uint64_t s = 0;
    for (int i = 0; i < SCAN_I; ++i) {
        uint64_t idx = base + (uint64_t)i * SCAN_T + threadIdx.x;
        if (idx < n) s += in[idx];
    }
    uint64_t tot;
    block_incl_scan(s, sh, &tot);
    if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

__global__ __launch_bounds__(SCAN_T) void k_scan_parts(uint64_t *part, uint32_t np)
{
    __shared__ uint64_t sh[SCAN_T / 64];
    uint64_t carry = 0;
    for (uint32_t b = 0; b < np; b += SCAN_T) {
        uint32_t i = b + threadIdx.x;
        uint64_t v = i < np ? part[i] : 0;
        uint64_t tot;
        uint64_t inc = block_incl_scan(v, sh, &tot);
        if (i < np) part[i] = carry + inc - v;  // exclusive
        carry += tot;
        __syncthreads();
    }
    if (threadIdx.x == 0) part[np] = carry;
}

__global__ __launch_bounds__(SCAN_T) void k_scan_final(const uint32_t *in, uint32_t *out, uint32_t n,
                                                        const uint64_t *part, uint32_t np)
{
    __shared__ uint64_t sh[SCAN_T / 64];
    // each thread owns SCAN_I consecutive items
    uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE + (uint64_t)threadIdx.x * SCAN_I;
    uint32_t v[SCAN_I];
    uint64_t s = 0;
#pragma unroll
    for (int i = 0; i < SCAN_I; ++i) {
        v[i] = (base + i < n) ? in[base + i] : 0;
        s += v[i];
    }
    uint64_t tot;
    uint64_t inc = block_incl_scan(s, sh, &tot);
    uint64_t run = part[blockIdx.x] + inc - s;
#pragma unroll
    for (int i = 0; i < SCAN_I; ++i) {
        if (base + i < n) out[base + i] = (uint32_t)run;
        run += v[i];
    }
    if (blockIdx.x == np - 1 && threadIdx.x == 0) out[n] = (uint32_t)part[np];
}

__global__ void k_fill_u32(uint32_t *a, uint32_t n, uint32_t v);

int scan_exclusive_u32(const uint32_t *in, uint32_t *out, uint32_t n, DevBuf &tmp, hipStream_t s)
{
    if (n == 0) {  // a fill kernel, not a memset node (the caller may be capturing a graph)
        hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, out, 1u, 0u);
        KC_HIP(hipGetLastError());
        return KC_OK;
    }
    uint32_t np = (n + SCAN_TILE - 1) / SCAN_TILE;
    int rc = tmp.ensure(sizeof(uint64_t) * (np + 1));
    if (rc) return rc;
    uint64_t *part = tmp.as<uint64_t>();
    hipLaunchKernelGGL(k_scan_reduce, dim3(np), dim3(SCAN_T), 0, s, in, n, part);
    hipLaunchKernelGGL(k_scan_parts, dim3(1), dim3(SCAN_T), 0, s, part, np);
    hipLaunchKernelGGL(k_scan_final, dim3(np), dim3(SCAN_T), 0, s, in, out, n, part, np);
    KC_HIP(hipGetLastError());
    return KC_OK;
}

// ---------------------------------------------------------------------------
// orientation
// ---------------------------------------------------------------------------
// keep(u->v) <=> (key u, u) < (key v, v)   (include/main_support.cuh:54-128)
__device__ inline bool keep_edge(uint32_t ku, uint32_t u, uint32_t kv, uint32_t v)
{
    return ku < kv || (ku == kv && u < v);
}

template <bool DEGEN>
__device__ inline uint32_t vkey(const uint32_t *rowptr, const uint32_t *key, uint32_t v)
{
    if (DEGEN) return key[v];
    return rowptr[v + 1] - rowptr[v];
}

// one wave per row: out-degree after orientation; also records the key
template <bool DEGEN>
__global__ __launch_bounds__(256) void k_orient_count(uint32_t n, const uint32_t *rowptr, const uint32_t *col,
                                                       uint32_t *key, uint32_t *outdeg)
{
    int lane = threadIdx.x & 63;
    uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    for (uint64_t u = wave; u < n; u += nw) {
        uint32_t b = rowptr[u], e = rowptr[u + 1];
        uint32_t ku = DEGEN ? key[u] : e - b;
        uint32_t cnt = 0;
        for (uint32_t i = b; i < e; i += 64) {
            bool k = false;
            if (i + lane < e) {
                uint32_t v = col[i + lane];
                k = keep_edge(ku, (uint32_t)u, vkey<DEGEN>(rowptr, key, v), v);
            }
            cnt += (uint32_t)__popcll(__ballot(k));
        }
        if (lane == 0) {
            outdeg[u] = cnt;
            if (!DEGEN) key[u] = ku;
        }
    }
}

// one wave per row: stable write of kept entries at ballot-prefix ranks
template <bool DEGEN>
__global__ __launch_bounds__(256) void k_orient_write(uint32_t n, const uint32_t *rowptr, const uint32_t *col,
                                                       const uint32_t *key, const uint32_t *orowptr,
                                                       uint32_t *ocol, uint32_t *osrc)
{
    int lane = threadIdx.x & 63;
    uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    for (uint64_t u = wave; u < n; u += nw) {
        uint32_t b = rowptr[u], e = rowptr[u + 1];
        uint32_t ku = DEGEN ? key[u] : e - b;
        uint32_t w = orowptr[u];
        for (uint32_t i = b; i < e; i += 64) {
            bool k = false;
            uint32_t v = 0;
            if (i + lane < e) {
                v = col[i + lane];
                k = keep_edge(ku, (uint32_t)u, vkey<DEGEN>(rowptr, key, v), v);
            }
            uint64_t bal = __ballot(k);
            if (k) {
                uint32_t r = w + (uint32_t)__popcll(bal & lt);
                ocol[r] = v;
                osrc[r] = (uint32_t)u;
            }
            w += (uint32_t)__popcll(bal);
        }
    }
}

__global__ void k_max_u32(const uint32_t *a, uint32_t n, uint32_t *out)
{
    uint32_t m = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        m = max(m, a[i]);
    for (int d = 32; d >= 1; d >>= 1) m = max(m, (uint32_t)__shfl_xor(m, d, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(out, m);
}

// ---------------------------------------------------------------------------
// batched degree orientation: many graphs concatenated, one launch per phase.
// Graph g owns rows [vbase[g], vbase[g+1]) of the concatenation; its full
// CSR is frow + rbase[g] (n_g + 1 local offsets) and fcol + ebase[g] (local
// ids); its oriented CSR goes to orow + rbase[g], and to ocol / osrc from
// scan[vbase[g]] on, where scan is the exclusive scan of all out-degrees.
// ---------------------------------------------------------------------------
struct MultiDesc {
    const uint32_t *vbase;  // [B+1] first row of each graph (global)
    const uint64_t *ebase;  // [B] first full-CSR entry of each graph
    uint32_t nb;            // graphs
    uint32_t nrows;         // total rows
};

__device__ inline uint32_t graph_of_row(const MultiDesc &d, uint32_t r)
{
    uint32_t lo = 0, hi = d.nb - 1;  // last g with vbase[g] <= r
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (d.vbase[mid] <= r) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__global__ __launch_bounds__(256) void k_orient_count_multi(MultiDesc d, const uint32_t *frow, const uint32_t *fcol,
                                                             uint32_t *key, uint32_t *outdeg)
{
    const int lane = threadIdx.x & 63;
    const uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    for (uint64_t r = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; r < d.nrows; r += nw) {
        const uint32_t g = graph_of_row(d, (uint32_t)r), u = (uint32_t)r - d.vbase[g];
        const uint32_t *rp = frow + d.vbase[g] + g, *col = fcol + d.ebase[g];
        const uint32_t b = rp[u], e = rp[u + 1], ku = e - b;
        uint32_t cnt = 0;
        for (uint32_t i = b; i < e; i += 64) {
            bool k = false;
            if (i + lane < e) {
                const uint32_t v = col[i + lane];
                k = keep_edge(ku, u, rp[v + 1] - rp[v], v);
            }
            cnt += (uint32_t)__popcll(__ballot(k));
        }
        if (lane == 0) {
            outdeg[r] = cnt;
            key[r] = ku;
        }
    }
}

__global__ __launch_bounds__(256) void k_orient_write_multi(MultiDesc d, const uint32_t *frow, const uint32_t *fcol,
                                                             const uint32_t *scan, uint32_t *orow, uint32_t *ocol,
                                                             uint32_t *osrc)
{
    const int lane = threadIdx.x & 63;
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    for (uint64_t r = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; r < d.nrows; r += nw) {
        const uint32_t g = graph_of_row(d, (uint32_t)r), u = (uint32_t)r - d.vbase[g];
        const uint32_t *rp = frow + d.vbase[g] + g, *col = fcol + d.ebase[g];
        const uint32_t b = rp[u], e = rp[u + 1], ku = e - b;
        const uint32_t gstart = scan[d.vbase[g]];
        uint32_t w = scan[r];
        if (lane == 0) {
            orow[d.vbase[g] + g + u] = w - gstart;
            if (r + 1 == d.vbase[g + 1]) orow[d.vbase[g] + g + u + 1] = scan[r + 1] - gstart;  // the row end
        }
        for (uint32_t i = b; i < e; i += 64) {
            bool k = false;
            uint32_t v = 0;
            if (i + lane < e) {
                v = col[i + lane];
                k = keep_edge(ku, u, rp[v + 1] - rp[v], v);
            }
            const uint64_t bal = __ballot(k);
            if (k) {
                const uint32_t at = w + (uint32_t)__popcll(bal & lt);
                ocol[at] = v;
                osrc[at] = u;
            }
            w += (uint32_t)__popcll(bal);
        }
    }
}

// per-graph max out-degree; a wave whose 64 rows lie in one graph does one atomic
__global__ __launch_bounds__(256) void k_max_multi(MultiDesc d, const uint32_t *outdeg, uint32_t *maxout)
{
    const int lane = threadIdx.x & 63;
    for (uint64_t base = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) & ~63ull; base < d.nrows;
         base += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t r = base + lane;
        const bool in = r < d.nrows;
        const uint32_t g = in ? graph_of_row(d, (uint32_t)r) : 0xFFFFFFFFu;
        const uint32_t v = in ? outdeg[r] : 0u;
        const uint32_t g0 = __shfl(g, 0, 64);
        if (__ballot(in && g != g0) == 0) {
            uint32_t m = v;
            for (int dd = 32; dd >= 1; dd >>= 1) m = max(m, (uint32_t)__shfl_xor(m, dd, 64));
            if (lane == 0) atomicMax(&maxout[g0], m);
        } else if (in) {
            atomicMax(&maxout[g], v);
        }
    }
}

__global__ void k_fill_u32(uint32_t *a, uint32_t n, uint32_t v)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        a[i] = v;
}

// ---------------------------------------------------------------------------
// k-core peel (kcore/kcore.cuh:442-584) as one cooperative kernel.
// Level L collects the live vertices of degree L; each round gives its
// frontier priority = round, then every frontier vertex decrements its live
// neighbours and a neighbour whose degree was L+1 joins the next round.
// The round that empties the graph breaks before assigning (kcore.cuh:495-498)
// so those vertices keep priority = nnz.  All blocks take the same decisions
// from counters read after a grid barrier.
// ---------------------------------------------------------------------------
struct PeelArgs {
    uint32_t n;
    uint32_t maxdeg;
    const uint32_t *rowptr;
    const uint32_t *col;
    int *deg;
    uint32_t *prio;
    uint8_t *state;  // 0 live, 1 peeled or in frontier
    uint32_t *q0, *q1;
    uint32_t *cnt;   // [2] queue counters, [2] max core, [3] error flag
};

__device__ inline void wave_append(bool take, uint32_t v, uint32_t *q, uint32_t *cnt)
{
    int lane = threadIdx.x & 63;
    uint64_t bal = __ballot(take);
    if (!bal) return;
    int leader = __ffsll((unsigned long long)bal) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(cnt, (uint32_t)__popcll(bal));
    base = __shfl(base, leader, 64);
    uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    if (take) q[base + (uint32_t)__popcll(bal & lt)] = v;
}

__global__ __launch_bounds__(256) void k_peel(PeelArgs a)
{
    cg::grid_group grid = cg::this_grid();
    const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t nthr = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t wave = tid >> 6, nwave = nthr >> 6;
    const int lane = threadIdx.x & 63;
    uint32_t *q[2] = {a.q0, a.q1};
    int c = 0;
    int64_t todo = a.n;
    uint32_t level = 0, round = 0;
    while (todo > 0) {
        // reset the gather counter, then gather live vertices of degree == level
        if (tid == 0) a.cnt[c] = 0;
        grid.sync();
        for (uint64_t base = wave * 64; base < a.n; base += nwave * 64) {
            uint64_t v = base + lane;
            bool take = v < a.n && a.state[v] == 0 && a.deg[v] == (int)level;
            wave_append(take, (uint32_t)v, q[c], &a.cnt[c]);
        }
        grid.sync();
        uint32_t nc = a.cnt[c];
        bool final = false;
        while (nc > 0) {
            todo -= nc;
            if (todo == 0) {
                final = true;
                break;
            }
            if (tid == 0) a.cnt[c ^ 1] = 0;
            for (uint64_t i = tid; i < nc; i += nthr) {
                uint32_t v = q[c][i];
                a.state[v] = 1;
                a.prio[v] = round;
            }
            grid.sync();
            // decrements: one wave per frontier vertex
            for (uint64_t i = wave; i < nc; i += nwave) {
                uint32_t v = q[c][i];
                uint32_t b = a.rowptr[v], e = a.rowptr[v + 1];
                for (uint32_t j = b; j < e; j += 64) {
                    bool take = false;
                    uint32_t x = 0;
                    if (j + lane < e) {
                        x = a.col[j + lane];
                        if (a.state[x] == 0) {
                            int old = atomicSub(&a.deg[x], 1);
                            take = (old == (int)level + 1);
                        }
                    }
                    wave_append(take, x, q[c ^ 1], &a.cnt[c ^ 1]);
                }
            }
            grid.sync();
            c ^= 1;
            ++round;
            nc = a.cnt[c];
        }
        if (final) {
            if (tid == 0) a.cnt[2] = level;
            break;
        }
        // the counter just read (cnt[c] == 0) is reset again by thread 0 at the
        // top of the next level; make every block finish reading it first
        grid.sync();
        ++level;
        if (level > a.maxdeg + 1) {
            if (tid == 0) a.cnt[3] = 1;
            break;
        }
    }
}

// ---------------------------------------------------------------------------
// The same peel as ordinary launches (the default path; it also serves
// processes that cannot get a cooperative launch).  A one-thread control
// kernel keeps the state machine that k_peel keeps in registers; the host
// enqueues steps of five kernels in batches and reads the phase once per
// batch.  Kernels of a step that has nothing to do return at once, so a
// batch that runs past the end is cheap.
// ---------------------------------------------------------------------------
struct PeelCtl {
    uint32_t phase;  // 0 gather, 1 round, 2 done
    uint32_t level, round, c;
    int64_t todo;
};

__global__ void k_peel_ctl_init(PeelCtl *ctl, uint32_t *cnt, uint32_t n)
{
    ctl->phase = n ? 0 : 2;
    ctl->level = ctl->round = ctl->c = 0;
    ctl->todo = n;
    cnt[0] = cnt[1] = cnt[2] = cnt[3] = 0;
}

__global__ __launch_bounds__(256) void k_peel_gather(PeelArgs a, const PeelCtl *ctl)
{
    if (ctl->phase != 0) return;
    const uint32_t c = ctl->c, level = ctl->level;
    uint32_t *q = c ? a.q1 : a.q0;
    const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t wave = tid >> 6, nwave = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    const int lane = threadIdx.x & 63;
    for (uint64_t base = wave * 64; base < a.n; base += nwave * 64) {
        uint64_t v = base + lane;
        bool take = v < a.n && a.state[v] == 0 && a.deg[v] == (int)level;
        wave_append(take, (uint32_t)v, q, &a.cnt[c]);
    }
}

// after a gather: an empty level moves on, a frontier starts a round, the
// frontier that empties the graph ends the peel unassigned
__global__ void k_peel_ctl_gathered(PeelCtl *ctl, uint32_t *cnt, uint32_t maxdeg)
{
    if (ctl->phase != 0) return;
    const uint32_t c = ctl->c, nc = cnt[c];
    if (nc == 0) {
        if (++ctl->level > maxdeg + 1) {
            cnt[3] = 1;
            ctl->phase = 2;
        }
        return;  // cnt[c] is still 0 for the next gather
    }
    ctl->todo -= nc;
    if (ctl->todo == 0) {
        cnt[2] = ctl->level;
        ctl->phase = 2;
        return;
    }
    cnt[c ^ 1] = 0;
    ctl->phase = 1;
}

__global__ __launch_bounds__(256) void k_peel_assign(PeelArgs a, const PeelCtl *ctl)
{
    if (ctl->phase != 1) return;
    const uint32_t c = ctl->c, nc = a.cnt[c], round = ctl->round;
    const uint32_t *q = c ? a.q1 : a.q0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nc; i += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t v = q[i];
        a.state[v] = 1;
        a.prio[v] = round;
    }
}

__global__ __launch_bounds__(256) void k_peel_decrement(PeelArgs a, const PeelCtl *ctl)
{
    if (ctl->phase != 1) return;
    const uint32_t c = ctl->c, nc = a.cnt[c], level = ctl->level;
    const uint32_t *q = c ? a.q1 : a.q0;
    uint32_t *qn = c ? a.q0 : a.q1;
    const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t wave = tid >> 6, nwave = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    const int lane = threadIdx.x & 63;
    for (uint64_t i = wave; i < nc; i += nwave) {
        uint32_t v = q[i];
        uint32_t b = a.rowptr[v], e = a.rowptr[v + 1];
        for (uint32_t j = b; j < e; j += 64) {
            bool take = false;
            uint32_t x = 0;
            if (j + lane < e) {
                x = a.col[j + lane];
                if (a.state[x] == 0) {
                    int old = atomicSub(&a.deg[x], 1);
                    take = (old == (int)level + 1);
                }
            }
            wave_append(take, x, qn, &a.cnt[c ^ 1]);
        }
    }
}

// after a round: the next frontier continues the level, an empty one ends it
__global__ void k_peel_ctl_round(PeelCtl *ctl, uint32_t *cnt, uint32_t maxdeg)
{
    if (ctl->phase != 1) return;
    const uint32_t c = ctl->c ^ 1;
    ctl->c = c;
    ctl->round++;
    const uint32_t nc = cnt[c];
    if (nc > 0) {
        ctl->todo -= nc;
        if (ctl->todo == 0) {
            cnt[2] = ctl->level;
            ctl->phase = 2;
            return;
        }
        cnt[c ^ 1] = 0;
        return;
    }
    if (++ctl->level > maxdeg + 1) {
        cnt[3] = 1;
        ctl->phase = 2;
        return;
    }
    cnt[c] = 0;
    ctl->phase = 0;
}

__global__ void k_peel_init(uint32_t n, const uint32_t *rowptr, int *deg, uint32_t *prio, uint8_t *state,
                            uint32_t nnz)
{
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (uint64_t)gridDim.x * blockDim.x) {
        deg[v] = (int)(rowptr[v + 1] - rowptr[v]);
        prio[v] = nnz;
        state[v] = 0;
    }
}

// A stream's launches captured once into a hipGraph and replayed.  When the
// runtime refuses the capture the same launches are enqueued on the stream
// each time instead (same work, one launch each).
struct Captured {
    hipGraph_t g = nullptr;
    hipGraphExec_t x = nullptr;
    ~Captured()
    {
        if (x) hipGraphExecDestroy(x);
        if (g) hipGraphDestroy(g);
    }
    template <class F>
    void capture(hipStream_t s, F &&enqueue)
    {
        if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) == hipSuccess) {
            enqueue();
            const hipError_t ce = hipStreamEndCapture(s, &g);
            if (ce != hipSuccess || !g || hipGraphInstantiateWithFlags(&x, g, 0) != hipSuccess) x = nullptr;
        }
        (void)hipGetLastError();
    }
    template <class F>
    hipError_t launch(hipStream_t s, F &&enqueue)
    {
        if (x) return hipGraphLaunch(x, s);
        enqueue();
        return hipGetLastError();
    }
};

static int run_peel(DeviceCtx *ctx, hipStream_t s, uint32_t n, uint32_t nnz, uint32_t maxdeg, const uint32_t *rowptr,
                    const uint32_t *col, uint32_t *prio, uint32_t *max_core)
{
    DevBuf deg, st, q0, q1, cnt;
    int rc;
    if ((rc = deg.ensure(sizeof(int) * (size_t)n + 4)) || (rc = st.ensure((size_t)n + 4)) ||
        (rc = q0.ensure(sizeof(uint32_t) * (size_t)n + 4)) || (rc = q1.ensure(sizeof(uint32_t) * (size_t)n + 4)) ||
        (rc = cnt.ensure(sizeof(uint32_t) * 4)))
        return rc;
    int gi = (int)std::min<uint64_t>(((uint64_t)n + 255) / 256, (uint64_t)ctx->cu * 8);
    if (gi < 1) gi = 1;
    hipLaunchKernelGGL(k_peel_init, dim3(gi), dim3(256), 0, s, n, rowptr, deg.as<int>(), prio, st.as<uint8_t>(), nnz);
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, cnt.as<uint32_t>(), 4u, 0u);  // counters, max core, error
    KC_HIP(hipGetLastError());
    int per_cu = 0;
    KC_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)k_peel, 256, 0));
    if (per_cu < 1) return fail(KC_ERR_DEVICE, "peel kernel cannot be resident");
    int grid = std::min(per_cu, 4) * ctx->cu;
    PeelArgs a;
    a.n = n;
    a.maxdeg = maxdeg;
    a.rowptr = rowptr;
    a.col = col;
    a.deg = deg.as<int>();
    a.prio = prio;
    a.state = st.as<uint8_t>();
    a.q0 = q0.as<uint32_t>();
    a.q1 = q1.as<uint32_t>();
    a.cnt = cnt.as<uint32_t>();
    void *args[] = {&a};
    // The stepped peel is the default: on the 1.7M-vertex power-law graph it
    // takes 93 ms against 384 ms for the cooperative kernel, whose three grid
    // barriers per round (1312 rounds there) dominate.  KC_PEEL_MODE=coop
    // selects the cooperative kernel (falling back when it is refused).
    const char *mode = std::getenv("KC_PEEL_MODE");
    bool coop = mode && std::strcmp(mode, "coop") == 0;
    if (coop && hipLaunchCooperativeKernel((const void *)k_peel, dim3(grid), dim3(256), args, 0, s) != hipSuccess) {
        (void)hipGetLastError();  // the launch was refused, nothing ran: step the peel from the host
        coop = false;
    }
    if (!coop) {
        // Steps of five launches (gather, its control, assign, decrement, the
        // round's control), captured once as a graph of PEEL_STEPS steps and
        // replayed until the control kernel reports the end: one graph launch
        // and one 4-byte read per PEEL_STEPS peel rounds or levels.  Steps past
        // the end return at once.
        constexpr int PEEL_STEPS = 64;
        DevBuf ctlb;
        if ((rc = ctlb.ensure(sizeof(PeelCtl)))) return rc;
        PeelCtl *ctl = ctlb.as<PeelCtl>();
        hipLaunchKernelGGL(k_peel_ctl_init, dim3(1), dim3(1), 0, s, ctl, a.cnt, n);
        KC_HIP(hipGetLastError());
        const int gs = std::max(1, std::min(gi, ctx->cu * 4));
        auto steps_enqueue = [&]() {
            for (int i = 0; i < PEEL_STEPS; ++i) {
                hipLaunchKernelGGL(k_peel_gather, dim3(gs), dim3(256), 0, s, a, ctl);
                hipLaunchKernelGGL(k_peel_ctl_gathered, dim3(1), dim3(1), 0, s, ctl, a.cnt, maxdeg);
                hipLaunchKernelGGL(k_peel_assign, dim3(gs), dim3(256), 0, s, a, ctl);
                hipLaunchKernelGGL(k_peel_decrement, dim3(gs), dim3(256), 0, s, a, ctl);
                hipLaunchKernelGGL(k_peel_ctl_round, dim3(1), dim3(1), 0, s, ctl, a.cnt, maxdeg);
            }
        };
        Captured steps_graph;
        steps_graph.capture(s, steps_enqueue);
        // every step peels a vertex or moves up a level: n + maxdeg + 2 steps end it
        const uint64_t max_steps = (uint64_t)n + maxdeg + 2;
        uint64_t steps = 0;
        uint32_t phase = 0;
        while (phase != 2) {
            if (steps > max_steps) return fail(KC_ERR_DEVICE, "peel did not converge (internal error)");
            KC_HIP(steps_graph.launch(s, steps_enqueue));
            steps += PEEL_STEPS;
            KC_HIP(hipMemcpyAsync(&phase, &ctl->phase, sizeof phase, hipMemcpyDeviceToHost, s));
            KC_HIP(hipStreamSynchronize(s));
        }
    }
    uint32_t h[4];
    KC_HIP(hipMemcpyAsync(h, cnt.p, sizeof h, hipMemcpyDeviceToHost, s));
    KC_HIP(hipStreamSynchronize(s));
    if (h[3]) return fail(KC_ERR_DEVICE, "peel did not converge (internal error)");
    *max_core = h[2];
    return KC_OK;
}

}  // namespace kc

using namespace kc;

kc_oriented::~kc_oriented() {}

extern "C" {

int kc_get_device_info(kc_device_info *info)
{
    if (!info) return fail(KC_ERR_ARG, "kc_get_device_info: null argument");
    DeviceCtx *ctx;
    int rc = device_ctx(&ctx);
    if (rc) return rc;
    std::memset(info, 0, sizeof *info);
    std::strncpy(info->name, ctx->prop.name, sizeof info->name - 1);
    std::strncpy(info->arch, ctx->prop.gcnArchName, sizeof info->arch - 1);
    info->cu_count = ctx->cu;
    info->hbm_bytes = ctx->hbm;
    info->lds_per_block = ctx->lds_per_block;
    info->clock_khz = ctx->prop.clockRate;
    return KC_OK;
}

int kc_select_device(int ordinal)
{
    int cnt = 0;
    if (hipGetDeviceCount(&cnt) != hipSuccess || cnt == 0) return fail(KC_ERR_DEVICE, "no HIP device");
    if (ordinal < 0 || ordinal >= cnt) return fail(KC_ERR_ARG, "device %d outside 0..%d", ordinal, cnt - 1);
    {
        std::lock_guard<std::mutex> lk(g_mu);
        if (hipSetDevice(ordinal) != hipSuccess) return fail(KC_ERR_DEVICE, "hipSetDevice(%d) failed", ordinal);
        if (g_state == 1 && g_ctx.dev != ordinal) {  // the old device keeps its streams
            g_ctx.stream = nullptr;
            for (hipStream_t &st : g_ctx.side) st = nullptr;
        }
        g_state = 0;  // re-read the properties of the selected device
    }
    DeviceCtx *ctx;
    return device_ctx(&ctx);
}

// Wait for and destroy the library's streams; the next call sets the device
// up again.  Run before the process exits so that no HIP call is left for
// the runtime's own teardown.
int kc_shutdown(void)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_state == 1) {
        hipSetDevice(g_ctx.dev);
        if (g_ctx.stream) {
            hipStreamSynchronize(g_ctx.stream);
            hipStreamDestroy(g_ctx.stream);
            g_ctx.stream = nullptr;
        }
        for (hipStream_t &st : g_ctx.side)
            if (st) {
                hipStreamSynchronize(st);
                hipStreamDestroy(st);
                st = nullptr;
            }
    }
    g_state = 0;
    return KC_OK;
}

int kc_orient(const kc_graph *g, int mode, kc_oriented **out)
{
    if (!g || !out) return fail(KC_ERR_ARG, "kc_orient: null argument");
    if (mode != KC_ORIENT_DEGREE && mode != KC_ORIENT_DEGEN) return fail(KC_ERR_ARG, "kc_orient: bad mode %d", mode);
    *out = nullptr;
    DeviceCtx *ctx;
    int rc = device_ctx(&ctx);
    if (rc) return rc;
    const uint32_t n = g->csr.n;
    const uint32_t nnz = (uint32_t)g->csr.col.size();
    hipStream_t s = ctx->stream;
    DevBuf frow, fcol;
    if ((rc = frow.ensure(sizeof(uint32_t) * ((size_t)n + 1))) || (rc = fcol.ensure(sizeof(uint32_t) * (size_t)nnz + 4)))
        return rc;
    if (hipMemcpyAsync(frow.p, g->csr.rowptr.data(), sizeof(uint32_t) * ((size_t)n + 1), hipMemcpyHostToDevice, s) !=
            hipSuccess ||
        (nnz && hipMemcpyAsync(fcol.p, g->csr.col.data(), sizeof(uint32_t) * (size_t)nnz, hipMemcpyHostToDevice, s) !=
                    hipSuccess))
        return fail(KC_ERR_DEVICE, "graph upload failed");
    return orient_device_csr(ctx, n, nnz, g->max_degree, frow.as<uint32_t>(), fcol.as<uint32_t>(), mode, out);
}

}  // extern "C"

// the orientation of a simple undirected CSR already in HBM (frow, fcol are
// only read; they may be freed once this returns)
int kc::orient_device_csr(DeviceCtx *ctx, uint32_t n, uint32_t nnz, uint32_t maxdeg, const uint32_t *frow_p,
                          const uint32_t *fcol_p, int mode, kc_oriented **out)
{
    *out = nullptr;
    int rc;
    // the CSR is simple and symmetric, so every undirected edge is kept once
    const uint32_t m_expect = nnz / 2;
    std::unique_ptr<kc_oriented> o(new kc_oriented);
    o->dev = ctx->dev;
    o->n = n;
    o->mode = mode;
    o->stream = ctx->stream;
    hipStream_t s = o->stream;
    // every buffer is sized up front, so the launches need no host read in
    // between and run as captured graphs
    DevBuf frow, fcol, outdeg, tmp, mx;
    frow.view(const_cast<uint32_t *>(frow_p), sizeof(uint32_t) * ((size_t)n + 1));
    fcol.view(const_cast<uint32_t *>(fcol_p), sizeof(uint32_t) * (size_t)nnz + 4);
    const size_t np = ((size_t)n + 2047) / 2048;
    if ((rc = outdeg.ensure(sizeof(uint32_t) * ((size_t)n + 1))) || (rc = o->key.ensure(sizeof(uint32_t) * (size_t)n + 4)) ||
        (rc = o->rowptr.ensure(sizeof(uint32_t) * ((size_t)n + 1))) ||
        (rc = o->col.ensure(sizeof(uint32_t) * (size_t)m_expect + 4)) ||
        (rc = o->src.ensure(sizeof(uint32_t) * (size_t)m_expect + 4)) || (rc = tmp.ensure(sizeof(uint64_t) * (np + 1))) ||
        (rc = mx.ensure(sizeof(uint32_t))))
        return rc;
    struct Events {
        hipEvent_t e0 = nullptr, e1 = nullptr;
        ~Events()
        {
            if (e0) hipEventDestroy(e0);
            if (e1) hipEventDestroy(e1);
        }
    } ev;
    KC_HIP(hipEventCreate(&ev.e0));
    KC_HIP(hipEventCreate(&ev.e1));
    KC_HIP(hipEventRecord(ev.e0, s));
    uint32_t max_core = 0;
    if (mode == KC_ORIENT_DEGEN && n > 0 &&
        (rc = run_peel(ctx, s, n, nnz, maxdeg, frow.as<uint32_t>(), fcol.as<uint32_t>(), o->key.as<uint32_t>(),
                       &max_core)))
        return rc;
    // keep flags, per-row counts, scan, stable compaction, max out-degree:
    // one captured graph
    int grid = (int)std::min<uint64_t>(((uint64_t)n + 3) / 4, (uint64_t)ctx->cu * 16);
    if (grid < 1) grid = 1;
    int scan_rc = KC_OK;
    auto enqueue = [&]() {
        const uint32_t *fr = frow.as<uint32_t>(), *fc = fcol.as<uint32_t>();
        uint32_t *key = o->key.as<uint32_t>(), *od = outdeg.as<uint32_t>(), *orow = o->rowptr.as<uint32_t>();
        if (n > 0) {
            if (mode == KC_ORIENT_DEGEN)
                hipLaunchKernelGGL(k_orient_count<true>, dim3(grid), dim3(256), 0, s, n, fr, fc, key, od);
            else
                hipLaunchKernelGGL(k_orient_count<false>, dim3(grid), dim3(256), 0, s, n, fr, fc, key, od);
        }
        scan_rc = scan_exclusive_u32(od, orow, n, tmp, s);
        if (n > 0) {
            if (mode == KC_ORIENT_DEGEN)
                hipLaunchKernelGGL(k_orient_write<true>, dim3(grid), dim3(256), 0, s, n, fr, fc, key, orow,
                                   o->col.as<uint32_t>(), o->src.as<uint32_t>());
            else
                hipLaunchKernelGGL(k_orient_write<false>, dim3(grid), dim3(256), 0, s, n, fr, fc, key, orow,
                                   o->col.as<uint32_t>(), o->src.as<uint32_t>());
        }
        hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, mx.as<uint32_t>(), 1u, 0u);
        if (n > 0)
            hipLaunchKernelGGL(k_max_u32, dim3(std::max(1, std::min(grid, 1024))), dim3(256), 0, s, od, n,
                               mx.as<uint32_t>());
    };
    Captured orient_graph;
    orient_graph.capture(s, enqueue);
    if (scan_rc) return scan_rc;
    KC_HIP(orient_graph.launch(s, enqueue));
    if (scan_rc) return scan_rc;
    KC_HIP(hipEventRecord(ev.e1, s));
    uint32_t res[2] = {0, 0};  // m, max out-degree
    KC_HIP(hipMemcpyAsync(&res[0], o->rowptr.as<uint32_t>() + n, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    KC_HIP(hipMemcpyAsync(&res[1], mx.p, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    const hipError_t se = hipStreamSynchronize(s);
    if (se != hipSuccess) return fail(KC_ERR_DEVICE, "orientation failed: %s", hipGetErrorString(se));
    if (res[0] != m_expect)
        return fail(KC_ERR_DEVICE, "orientation kept %u of %u edges (internal error)", res[0], m_expect);
    float ms = 0;
    hipEventElapsedTime(&ms, ev.e0, ev.e1);
    o->m = res[0];
    o->max_out = res[1];
    o->max_core = max_core;
    o->pre_ms = ms;
    *out = o.release();
    return KC_OK;
}

extern "C" {

// Build the CSR of a raw edge list on the device and orient it there: the
// host never holds the CSR.
int kc_orient_edges(uint32_t n, size_t nedges, const uint32_t *src, const uint32_t *dst, int mode, kc_oriented **out)
{
    if (!out || (nedges && (!src || !dst))) return fail(KC_ERR_ARG, "kc_orient_edges: null argument");
    if (mode != KC_ORIENT_DEGREE && mode != KC_ORIENT_DEGEN) return fail(KC_ERR_ARG, "kc_orient_edges: bad mode %d", mode);
    *out = nullptr;
    DeviceCtx *ctx;
    int rc = device_ctx(&ctx);
    if (rc) return rc;
    DevBuf rp, col;
    uint32_t nv = 0, maxdeg = 0;
    uint64_t nnz = 0;
    if ((rc = build_csr_device(ctx, ctx->stream, n, nedges, src, dst, rp, col, nv, nnz, maxdeg))) return rc;
    rc = orient_device_csr(ctx, nv, (uint32_t)nnz, maxdeg, rp.as<uint32_t>(), col.as<uint32_t>(), mode, out);
    hipStreamSynchronize(ctx->stream);  // rp and col are freed on return
    return rc;
}

// kc_graph_from_edges with the CSR built on the device and copied back
int kc_graph_from_edges_device(uint32_t n, size_t nedges, const uint32_t *src, const uint32_t *dst, kc_graph **out)
{
    if (!out || (nedges && (!src || !dst))) return fail(KC_ERR_ARG, "kc_graph_from_edges_device: null argument");
    *out = nullptr;
    DeviceCtx *ctx;
    int rc = device_ctx(&ctx);
    if (rc) return rc;
    DevBuf rp, col;
    uint32_t nv = 0, maxdeg = 0;
    uint64_t nnz = 0;
    if ((rc = build_csr_device(ctx, ctx->stream, n, nedges, src, dst, rp, col, nv, nnz, maxdeg))) return rc;
    std::unique_ptr<kc_graph> g(new kc_graph);
    g->csr.n = nv;
    g->csr.rowptr.resize((size_t)nv + 1);
    g->csr.col.resize(nnz);
    KC_HIP(hipMemcpyAsync(g->csr.rowptr.data(), rp.p, 4 * ((size_t)nv + 1), hipMemcpyDeviceToHost, ctx->stream));
    if (nnz) KC_HIP(hipMemcpyAsync(g->csr.col.data(), col.p, 4 * nnz, hipMemcpyDeviceToHost, ctx->stream));
    KC_HIP(hipStreamSynchronize(ctx->stream));
    g->max_degree = maxdeg;
    *out = g.release();
    return KC_OK;
}

// Batched orientation.  Degree-order graphs are concatenated into one pinned
// upload and oriented by one captured hipGraph (max-out reset, row counts,
// three scan kernels, compaction, per-graph max out-degree); their oriented
// CSRs are views into one shared allocation.  Degeneracy-order and empty
// graphs go through kc_orient one by one.
static int orient_degree_chunk(DeviceCtx *ctx, const kc_graph *const *gs, const std::vector<size_t> &idx,
                               kc_oriented **out)
{
    const size_t B = idx.size();
    hipStream_t s = ctx->stream;
    std::vector<uint32_t> vbase(B + 1, 0);
    std::vector<uint64_t> ebase(B, 0), cbase(B + 1, 0);
    uint64_t nent = 0;
    for (size_t j = 0; j < B; ++j) {
        const kc::HostCSR &c = gs[idx[j]]->csr;
        vbase[j + 1] = vbase[j] + c.n;
        ebase[j] = nent;
        nent += c.col.size();
        cbase[j + 1] = cbase[j] + c.col.size() / 2;
    }
    const uint32_t nrows = vbase[B];
    const uint64_t rp_words = (uint64_t)nrows + B, mtot = cbase[B];
    // pinned staging: full rowptrs, full cols, vbase, ebase
    const size_t stage_bytes = 4 * (rp_words + nent) + 4 * (B + 1) + 8 * B + 64;
    uint8_t *stage = nullptr;
    if (hipHostMalloc((void **)&stage, stage_bytes, hipHostMallocDefault) != hipSuccess)
        return fail(KC_ERR_NOMEM, "hipHostMalloc(%zu) failed", stage_bytes);
    uint32_t *h_frow = reinterpret_cast<uint32_t *>(stage);
    uint32_t *h_fcol = h_frow + rp_words;
    uint64_t *h_ebase = reinterpret_cast<uint64_t *>(stage + ((4 * (rp_words + nent) + 7) & ~(size_t)7));
    uint32_t *h_vbase = reinterpret_cast<uint32_t *>(h_ebase + B);
    {
        const unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        std::vector<std::thread> th;
        for (unsigned t = 0; t < nt; ++t)
            th.emplace_back([&, t] {
                for (size_t j = t; j < B; j += nt) {
                    const kc::HostCSR &c = gs[idx[j]]->csr;
                    std::memcpy(h_frow + vbase[j] + j, c.rowptr.data(), 4 * ((size_t)c.n + 1));
                    if (!c.col.empty()) std::memcpy(h_fcol + ebase[j], c.col.data(), 4 * c.col.size());
                }
            });
        for (auto &x : th) x.join();
    }
    std::memcpy(h_ebase, ebase.data(), 8 * B);
    std::memcpy(h_vbase, vbase.data(), 4 * (B + 1));

    // device: outputs in one shared arena, inputs and temporaries freed at the end
    auto arena = std::make_shared<DevBuf>();
    const uint64_t a_rows = rp_words, a_col = a_rows, a_src = a_col + mtot, a_key = a_src + mtot,
                   a_words = a_key + nrows + 4;
    DevBuf din, dtmp, dscan, dmeta;
    int rc;
    const size_t np = ((size_t)nrows + 2047) / 2048;
    if ((rc = arena->ensure(4 * a_words)) || (rc = din.ensure(4 * (rp_words + nent) + 8)) ||
        (rc = dscan.ensure(4 * (2 * (size_t)nrows + 2))) || (rc = dtmp.ensure(8 * (np + 1))) ||
        (rc = dmeta.ensure(8 * B + 4 * (B + 1) + 4 * B + 64))) {
        hipHostFree(stage);
        return rc;
    }
    uint32_t *d_orow = arena->as<uint32_t>(), *d_ocol = d_orow + a_col, *d_osrc = d_orow + a_src,
             *d_key = d_orow + a_key;
    uint32_t *d_frow = din.as<uint32_t>(), *d_fcol = d_frow + rp_words;
    uint32_t *d_outdeg = dscan.as<uint32_t>(), *d_scan = d_outdeg + nrows + 1;
    uint64_t *d_ebase = dmeta.as<uint64_t>();
    uint32_t *d_vbase = reinterpret_cast<uint32_t *>(d_ebase + B), *d_maxout = d_vbase + B + 1;
    KC_HIP(hipMemcpyAsync(d_frow, h_frow, 4 * (rp_words + nent), hipMemcpyHostToDevice, s));
    KC_HIP(hipMemcpyAsync(d_ebase, h_ebase, 8 * B, hipMemcpyHostToDevice, s));
    KC_HIP(hipMemcpyAsync(d_vbase, h_vbase, 4 * (B + 1), hipMemcpyHostToDevice, s));

    MultiDesc md{d_vbase, d_ebase, (uint32_t)B, nrows};
    const int grid = (int)std::max<uint64_t>(1, std::min<uint64_t>(((uint64_t)nrows + 3) / 4, (uint64_t)ctx->cu * 16));
    const int gmax = std::max(1, std::min(grid, 1024));
    auto enqueue = [&]() {
        hipLaunchKernelGGL(k_fill_u32, dim3(std::max(1, std::min((int)((B + 255) / 256), 1024))), dim3(256), 0, s,
                           d_maxout, (uint32_t)B, 0u);
        hipLaunchKernelGGL(k_orient_count_multi, dim3(grid), dim3(256), 0, s, md, d_frow, d_fcol, d_key, d_outdeg);
        const int r = scan_exclusive_u32(d_outdeg, d_scan, nrows, dtmp, s);
        hipLaunchKernelGGL(k_orient_write_multi, dim3(grid), dim3(256), 0, s, md, d_frow, d_fcol, d_scan, d_orow,
                           d_ocol, d_osrc);
        hipLaunchKernelGGL(k_max_multi, dim3(gmax), dim3(256), 0, s, md, d_outdeg, d_maxout);
        return r;
    };
    // one captured graph; if this runtime refuses the capture, the same launches go on the stream
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t e0, e1;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    bool captured = false;
    if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) == hipSuccess) {
        const int r = enqueue();
        const hipError_t ce = hipStreamEndCapture(s, &graph);
        captured = (r == KC_OK && ce == hipSuccess && graph &&
                    hipGraphInstantiateWithFlags(&exec, graph, 0) == hipSuccess);
    }
    (void)hipGetLastError();
    hipEventRecord(e0, s);
    hipError_t le = hipSuccess;
    if (captured) le = hipGraphLaunch(exec, s);
    else if ((rc = enqueue())) le = hipErrorUnknown;
    hipEventRecord(e1, s);
    std::vector<uint32_t> maxout(B);
    if (le == hipSuccess) le = hipMemcpyAsync(maxout.data(), d_maxout, 4 * B, hipMemcpyDeviceToHost, s);
    const hipError_t se = hipStreamSynchronize(s);
    float ms = 0;
    hipEventElapsedTime(&ms, e0, e1);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (exec) hipGraphExecDestroy(exec);
    if (graph) hipGraphDestroy(graph);
    hipHostFree(stage);
    if (rc) return rc;
    if (le != hipSuccess || se != hipSuccess)
        return fail(KC_ERR_DEVICE, "batched orientation failed: %s", hipGetErrorString(le != hipSuccess ? le : se));
    for (size_t j = 0; j < B; ++j) {
        const kc::HostCSR &c = gs[idx[j]]->csr;
        kc_oriented *o = new kc_oriented;
        o->dev = ctx->dev;
        o->stream = s;
        o->n = c.n;
        o->m = cbase[j + 1] - cbase[j];
        o->mode = KC_ORIENT_DEGREE;
        o->max_out = maxout[j];
        o->max_core = 0;
        o->pre_ms = ms;  // the whole batch's device time
        o->rowptr.view(d_orow + vbase[j] + j, 4 * ((size_t)c.n + 1));
        o->col.view(d_ocol + cbase[j], 4 * o->m);
        o->src.view(d_osrc + cbase[j], 4 * o->m);
        o->key.view(d_key + vbase[j], 4 * (size_t)c.n);
        o->arena = arena;
        out[idx[j]] = o;
    }
    return KC_OK;
}

extern "C" int kc_orient_batch(const kc_graph *const *gs, const int *modes, size_t nb, kc_oriented **out)
{
    if (nb && (!gs || !modes || !out)) return fail(KC_ERR_ARG, "kc_orient_batch: null argument");
    for (size_t i = 0; i < nb; ++i) {
        out[i] = nullptr;
        if (!gs[i]) return fail(KC_ERR_ARG, "kc_orient_batch: null graph %zu", i);
        if (modes[i] != KC_ORIENT_DEGREE && modes[i] != KC_ORIENT_DEGEN)
            return fail(KC_ERR_ARG, "kc_orient_batch: bad mode %d for graph %zu", modes[i], i);
    }
    DeviceCtx *ctx;
    int rc = device_ctx(&ctx);
    if (rc) return rc;
    auto cleanup = [&](int code) {
        for (size_t i = 0; i < nb; ++i) {
            delete out[i];
            out[i] = nullptr;
        }
        return code;
    };
    // chunks of at most 2^31 rows + entries each (32-bit offsets inside a chunk)
    std::vector<size_t> chunk;
    uint64_t words = 0;
    for (size_t i = 0; i <= nb; ++i) {
        const bool end = (i == nb);
        uint64_t w = 0;
        if (!end) {
            const kc::HostCSR &c = gs[i]->csr;
            if (modes[i] == KC_ORIENT_DEGEN || c.n == 0) {
                if ((rc = kc_orient(gs[i], modes[i], &out[i]))) return cleanup(rc);
                continue;
            }
            w = (uint64_t)c.n + 1 + c.col.size();
        }
        if (!chunk.empty() && (end || words + w >= (1ull << 31))) {
            if ((rc = orient_degree_chunk(ctx, gs, chunk, out))) return cleanup(rc);
            chunk.clear();
            words = 0;
        }
        if (!end) {
            chunk.push_back(i);
            words += w;
        }
    }
    return KC_OK;
}

int kc_oriented_get_info(const kc_oriented *o, kc_oriented_info *info)
{
    if (!o || !info) return fail(KC_ERR_ARG, "kc_oriented_get_info: null argument");
    info->n = o->n;
    info->m = o->m;
    info->max_out_degree = o->max_out;
    info->max_core = o->max_core;
    info->preprocess_ms = o->pre_ms;
    return KC_OK;
}

int kc_oriented_export(const kc_oriented *o, uint32_t *rowptr, uint32_t *col, uint32_t *key)
{
    if (!o) return fail(KC_ERR_ARG, "kc_oriented_export: null handle");
    hipSetDevice(o->dev);
    if (rowptr) KC_HIP(hipMemcpy(rowptr, o->rowptr.p, sizeof(uint32_t) * ((size_t)o->n + 1), hipMemcpyDeviceToHost));
    if (col && o->m) KC_HIP(hipMemcpy(col, o->col.p, sizeof(uint32_t) * (size_t)o->m, hipMemcpyDeviceToHost));
    if (key && o->n) KC_HIP(hipMemcpy(key, o->key.p, sizeof(uint32_t) * (size_t)o->n, hipMemcpyDeviceToHost));
    return KC_OK;
}

void kc_oriented_free(kc_oriented *o) { delete o; }

}  // extern "C"

// ---------------------------------------------------------------------------
// binary oriented-CSR cache
// ---------------------------------------------------------------------------
namespace {

const char kCacheMagic[8] = {'K', 'C', 'O', 'C', 'S', 'R', '1', '\0'};

struct CacheHeader {
    char magic[8];
    uint32_t n, mode;
    uint64_t m;
    uint32_t max_out, max_core;
    uint64_t checksum;
};
static_assert(sizeof(CacheHeader) == 40, "cache header is 40 bytes");

uint64_t fnv1a_words(uint64_t h, const uint32_t *p, size_t count)
{
    for (size_t i = 0; i < count; ++i)
        for (int b = 0; b < 4; ++b) {
            h ^= (p[i] >> (8 * b)) & 0xFFu;
            h *= 1099511628211ull;
        }
    return h;
}

uint64_t cache_checksum(const std::vector<uint32_t> &rp, const std::vector<uint32_t> &col,
                        const std::vector<uint32_t> &key)
{
    uint64_t h = 14695981039346656037ull;
    h = fnv1a_words(h, rp.data(), rp.size());
    h = fnv1a_words(h, col.data(), col.size());
    return fnv1a_words(h, key.data(), key.size());
}

}  // namespace

extern "C" int kc_oriented_save(const kc_oriented *o, const char *path)
{
    if (!o || !path) return fail(KC_ERR_ARG, "kc_oriented_save: null argument");
    std::vector<uint32_t> rp((size_t)o->n + 1), col(o->m), key(o->n);
    int rc = kc_oriented_export(o, rp.data(), col.data(), key.data());
    if (rc) return rc;
    CacheHeader h;
    std::memcpy(h.magic, kCacheMagic, 8);
    h.n = o->n;
    h.mode = (uint32_t)o->mode;
    h.m = o->m;
    h.max_out = o->max_out;
    h.max_core = o->max_core;
    h.checksum = cache_checksum(rp, col, key);
    FILE *f = std::fopen(path, "wb");
    if (!f) return fail(KC_ERR_IO, "cannot create '%s'", path);
    bool ok = std::fwrite(&h, sizeof h, 1, f) == 1 && std::fwrite(rp.data(), 4, rp.size(), f) == rp.size() &&
              std::fwrite(col.data(), 4, col.size(), f) == col.size() &&
              std::fwrite(key.data(), 4, key.size(), f) == key.size();
    if (std::fclose(f) != 0) ok = false;
    return ok ? KC_OK : fail(KC_ERR_IO, "write to '%s' failed", path);
}

extern "C" int kc_oriented_load(const char *path, kc_oriented **out)
{
    if (!path || !out) return fail(KC_ERR_ARG, "kc_oriented_load: null argument");
    *out = nullptr;
    FILE *f = std::fopen(path, "rb");
    if (!f) return fail(KC_ERR_IO, "cannot open '%s'", path);
    CacheHeader h;
    std::vector<uint32_t> rp, col, key;
    bool ok = std::fread(&h, sizeof h, 1, f) == 1 && std::memcmp(h.magic, kCacheMagic, 8) == 0 &&
              (h.mode == KC_ORIENT_DEGREE || h.mode == KC_ORIENT_DEGEN) && h.m < 0xFFFFFFFFull && h.n < 0xFFFFFFFFu;
    if (ok) {
        rp.resize((size_t)h.n + 1);
        col.resize(h.m);
        key.resize(h.n);
        ok = std::fread(rp.data(), 4, rp.size(), f) == rp.size() && std::fread(col.data(), 4, col.size(), f) == col.size() &&
             std::fread(key.data(), 4, key.size(), f) == key.size() && std::fgetc(f) == EOF;
    }
    std::fclose(f);
    if (!ok) return fail(KC_ERR_FORMAT, "'%s' is not an oriented-CSR cache file", path);
    if (cache_checksum(rp, col, key) != h.checksum) return fail(KC_ERR_FORMAT, "'%s': checksum mismatch", path);
    // well formed: monotone rowptr ending at m, sorted rows of ids < n, and
    // every arc going up in (key, id) order (so the graph is acyclic)
    uint32_t max_out = 0;
    if (rp[0] != 0 || rp[h.n] != h.m) return fail(KC_ERR_FORMAT, "'%s': bad row pointers", path);
    for (uint32_t u = 0; u < h.n; ++u) {
        if (rp[u + 1] < rp[u]) return fail(KC_ERR_FORMAT, "'%s': bad row pointers", path);
        max_out = std::max(max_out, rp[u + 1] - rp[u]);
        for (uint32_t i = rp[u]; i < rp[u + 1]; ++i) {
            uint32_t v = col[i];
            if (v >= h.n || (i > rp[u] && col[i - 1] >= v) || key[u] > key[v] || (key[u] == key[v] && u >= v))
                return fail(KC_ERR_FORMAT, "'%s': row %u is not a sorted upward out-list", path, u);
        }
    }
    if (max_out != h.max_out) return fail(KC_ERR_FORMAT, "'%s': max out-degree does not match the rows", path);
    DeviceCtx *ctx;
    int rc = device_ctx(&ctx);
    if (rc) return rc;
    kc_oriented *o = new kc_oriented;
    o->dev = ctx->dev;
    o->stream = ctx->stream;
    o->n = h.n;
    o->m = h.m;
    o->mode = (int)h.mode;
    o->max_out = max_out;
    o->max_core = h.max_core;
    std::vector<uint32_t> src(h.m);
    for (uint32_t u = 0; u < h.n; ++u)
        for (uint32_t i = rp[u]; i < rp[u + 1]; ++i) src[i] = u;
    if ((rc = o->rowptr.ensure(4 * ((size_t)h.n + 1))) || (rc = o->col.ensure(4 * (size_t)h.m + 4)) ||
        (rc = o->src.ensure(4 * (size_t)h.m + 4)) || (rc = o->key.ensure(4 * (size_t)h.n + 4))) {
        delete o;
        return rc;
    }
    hipStream_t s = o->stream;
    bool up = hipMemcpyAsync(o->rowptr.p, rp.data(), 4 * rp.size(), hipMemcpyHostToDevice, s) == hipSuccess &&
              (h.m == 0 || (hipMemcpyAsync(o->col.p, col.data(), 4 * col.size(), hipMemcpyHostToDevice, s) == hipSuccess &&
                            hipMemcpyAsync(o->src.p, src.data(), 4 * src.size(), hipMemcpyHostToDevice, s) == hipSuccess)) &&
              (h.n == 0 || hipMemcpyAsync(o->key.p, key.data(), 4 * key.size(), hipMemcpyHostToDevice, s) == hipSuccess) &&
              hipStreamSynchronize(s) == hipSuccess;
    if (!up) {
        delete o;
        return fail(KC_ERR_DEVICE, "oriented-CSR upload failed");
    }
    *out = o;
    return KC_OK;
}
