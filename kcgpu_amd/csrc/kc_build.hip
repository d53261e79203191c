// kc_build.hip -- the simple undirected CSR built on the device from a raw
// edge list: self-loops dropped, every edge stored in both directions, rows
// sorted, duplicates removed.  The same graph as the host builder (kc_io.cpp,
// build_csr), which replaces CSRCOO::from_edgelist (include/CSRCOO.cuh:37-80)
// plus the reference's optional std::sort (include/FIleReader.cuh:314-320).
//
// Steps: degree histogram (both endpoints of every non-loop edge), exclusive
// scan into row offsets, scatter of both directions through per-row atomic
// cursors (unordered within a row), a per-row sort that also removes
// duplicates, a scan of the unique counts and a compaction into the final
// CSR.  Rows of at most SMALL_ROW entries are sorted by one wave each with a
// bitonic network in LDS; longer rows (hubs of power-law graphs, up to
// hundreds of thousands of entries) by one block each through a bitmap over
// the vertex range in HBM: set a bit per entry, then read the set bits back
// in order, which sorts and deduplicates at once.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "kc_device.h"

namespace kc {

constexpr uint32_t SMALL_ROW = 2048;  // entries a wave sorts in LDS (8 KB)
constexpr int BUILD_T = 256;          // threads per block (4 waves)

__device__ inline uint32_t build_wave_max(uint32_t x)
{
    for (int d = 32; d >= 1; d >>= 1) x = max(x, (uint32_t)__shfl_xor(x, d, 64));
    return x;
}

// largest id over the non-loop edges, and whether there is one
__global__ __launch_bounds__(BUILD_T) void k_edge_max(const uint32_t *s, const uint32_t *d, uint64_t ne, uint32_t *mx,
                                                      uint32_t *any)
{
    uint32_t m = 0, a = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < ne; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t u = s[i], v = d[i];
        if (u == v) continue;
        a = 1;
        m = max(m, max(u, v));
    }
    m = build_wave_max(m);
    a = build_wave_max(a);
    if ((threadIdx.x & 63) == 0) {
        if (a) atomicOr(any, 1u);
        atomicMax(mx, m);
    }
}

__global__ __launch_bounds__(BUILD_T) void k_edge_degree(const uint32_t *s, const uint32_t *d, uint64_t ne,
                                                         uint32_t *deg)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < ne; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t u = s[i], v = d[i];
        if (u == v) continue;
        atomicAdd(&deg[u], 1u);
        atomicAdd(&deg[v], 1u);
    }
}

__global__ __launch_bounds__(BUILD_T) void k_edge_scatter(const uint32_t *s, const uint32_t *d, uint64_t ne,
                                                          uint32_t *cursor, uint32_t *tmp)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < ne; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t u = s[i], v = d[i];
        if (u == v) continue;
        tmp[atomicAdd(&cursor[u], 1u)] = v;
        tmp[atomicAdd(&cursor[v], 1u)] = u;
    }
}

// Rows of at most SMALL_ROW entries, one wave each: bitonic sort in LDS (the
// row padded with 0xFFFFFFFF to a power of two), then the distinct values
// written back to the row's start in order; ucnt[r] = how many.  Longer rows
// are listed in big[] for k_row_sort_big.
__global__ __launch_bounds__(BUILD_T) void k_row_sort_small(const uint32_t *off, uint32_t n, uint32_t *tmp,
                                                            uint32_t *ucnt, uint32_t *big, uint32_t *nbig)
{
    __shared__ uint32_t buf[BUILD_T / 64][SMALL_ROW];
    const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    uint32_t *a = buf[wv];
    const uint32_t nwave = gridDim.x * (BUILD_T / 64);
    for (uint32_t r = blockIdx.x * (BUILD_T / 64) + wv; r < n; r += nwave) {
        const uint32_t b = off[r], dn = off[r + 1] - b;
        if (dn > SMALL_ROW) {
            if (lane == 0) big[atomicAdd(nbig, 1u)] = r;
            continue;
        }
        if (dn <= 1) {
            if (lane == 0) ucnt[r] = dn;
            continue;
        }
        uint32_t p = 2;
        while (p < dn) p <<= 1;
        for (uint32_t i = lane; i < p; i += 64) a[i] = i < dn ? tmp[b + i] : 0xFFFFFFFFu;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        for (uint32_t kk = 2; kk <= p; kk <<= 1) {
            for (uint32_t j = kk >> 1; j > 0; j >>= 1) {
                for (uint32_t i = lane; i < p; i += 64) {
                    const uint32_t x = i ^ j;
                    if (x > i) {
                        const uint32_t ai = a[i], ax = a[x];
                        const bool up = (i & kk) == 0;
                        if ((ai > ax) == up) {
                            a[i] = ax;
                            a[x] = ai;
                        }
                    }
                }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            }
        }
        // distinct values, in order, to the row's start
        uint32_t pos = 0;
        for (uint32_t i0 = 0; i0 < dn; i0 += 64) {
            const uint32_t i = i0 + lane;
            const bool keep = i < dn && (i == 0 || a[i] != a[i - 1]);
            const uint64_t bal = __ballot(keep);
            if (keep) tmp[b + pos + (uint32_t)__popcll(bal & lt)] = a[i];
            pos += (uint32_t)__popcll(bal);
        }
        if (lane == 0) ucnt[r] = pos;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();  // a[] is reused by the next row
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

// Rows longer than SMALL_ROW, one block each: a bitmap of `words` 64-bit
// words (the vertex range) in this block's HBM slice; set a bit per entry,
// then emit the set bits in order (each thread a contiguous run of words,
// placed by a block-wide exclusive scan of the runs' bit counts).
__global__ __launch_bounds__(BUILD_T) void k_row_sort_big(const uint32_t *off, const uint32_t *big, uint32_t nbig,
                                                          uint32_t *tmp, uint32_t *ucnt, uint64_t *bitmaps,
                                                          uint32_t words)
{
    __shared__ uint32_t wsum[BUILD_T / 64];
    uint64_t *bm = bitmaps + (uint64_t)blockIdx.x * words;
    const uint32_t t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const uint32_t per = (words + BUILD_T - 1) / BUILD_T;  // words per thread
    for (uint32_t bi = blockIdx.x; bi < nbig; bi += gridDim.x) {
        const uint32_t r = big[bi], b = off[r], dn = off[r + 1] - b;
        for (uint32_t x = t; x < words; x += BUILD_T) bm[x] = 0;
        __syncthreads();
        for (uint32_t i = t; i < dn; i += BUILD_T) {
            const uint32_t v = tmp[b + i];
            atomicOr((unsigned long long *)&bm[v >> 6], 1ull << (v & 63));
        }
        __syncthreads();
        // (the bits were set by atomics at the device's coherence point: read
        // them there, not from this CU's cache)
        auto word = [&](uint32_t x) { return __hip_atomic_load(&bm[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); };
        const uint32_t w0 = min(words, t * per), w1 = min(words, w0 + per);
        uint32_t c = 0;
        for (uint32_t x = w0; x < w1; ++x) c += (uint32_t)__popcll(word(x));
        // block-wide exclusive scan of c
        uint32_t incl = c;
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t y = __shfl_up(incl, (unsigned)d, 64);
            if ((int)lane >= d) incl += y;
        }
        if (lane == 63) wsum[wv] = incl;
        __syncthreads();
        uint32_t before = incl - c, total = 0;
        for (uint32_t q = 0; q < BUILD_T / 64; ++q) {
            if (q < wv) before += wsum[q];
            total += wsum[q];
        }
        uint32_t pos = b + before;
        for (uint32_t x = w0; x < w1; ++x)
            for (uint64_t bits = word(x); bits; bits &= bits - 1) tmp[pos++] = x * 64u + (uint32_t)__builtin_ctzll(bits);
        if (t == 0) ucnt[r] = total;
        __syncthreads();  // bm and wsum are reused by the next row
    }
}

// final CSR: the first ucnt[r] entries of row r's segment, one wave a row
__global__ __launch_bounds__(BUILD_T) void k_row_compact(const uint32_t *off, const uint32_t *tmp,
                                                         const uint32_t *rowptr, uint32_t n, uint32_t *col)
{
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t nwave = gridDim.x * (BUILD_T / 64);
    for (uint32_t r = blockIdx.x * (BUILD_T / 64) + (threadIdx.x >> 6); r < n; r += nwave) {
        const uint32_t b = off[r], o = rowptr[r], c = rowptr[r + 1] - o;
        for (uint32_t i = lane; i < c; i += 64) col[o + i] = tmp[b + i];
    }
}

__global__ __launch_bounds__(BUILD_T) void k_build_max(const uint32_t *a, uint32_t n, uint32_t *out)
{
    uint32_t m = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) m = max(m, a[i]);
    m = build_wave_max(m);
    if ((threadIdx.x & 63) == 0) atomicMax(out, m);
}

__global__ void k_build_fill(uint32_t *a, uint64_t n, uint32_t v)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        a[i] = v;
}

int build_csr_device(DeviceCtx *ctx, hipStream_t s, uint32_t n_in, size_t ne, const uint32_t *hsrc,
                     const uint32_t *hdst, DevBuf &rowptr, DevBuf &col, uint32_t &n_out, uint64_t &nnz_out,
                     uint32_t &maxdeg_out)
{
    if (2.0 * (double)ne >= 4294967295.0)
        return fail(KC_ERR_LIMIT, "%zu edges: %zu CSR entries exceed 32-bit offsets", ne, 2 * ne);
    DevBuf es, ed, small;
    int rc;
    if ((rc = es.ensure(4 * ne + 4)) || (rc = ed.ensure(4 * ne + 4)) || (rc = small.ensure(64))) return rc;
    if (ne) {
        KC_HIP(hipMemcpyAsync(es.p, hsrc, 4 * ne, hipMemcpyHostToDevice, s));
        KC_HIP(hipMemcpyAsync(ed.p, hdst, 4 * ne, hipMemcpyHostToDevice, s));
    }
    const uint32_t *S = es.as<uint32_t>(), *D = ed.as<uint32_t>();
    const int ge = (int)std::max<uint64_t>(1, std::min<uint64_t>((ne + BUILD_T - 1) / BUILD_T, (uint64_t)ctx->cu * 16));
    uint32_t *sm = small.as<uint32_t>();  // [0] max id, [1] any edge, [2] big rows, [3] max degree
    hipLaunchKernelGGL(k_build_fill, dim3(1), dim3(64), 0, s, sm, (uint64_t)16, 0u);
    if (ne) hipLaunchKernelGGL(k_edge_max, dim3(ge), dim3(BUILD_T), 0, s, S, D, (uint64_t)ne, sm, sm + 1);
    KC_HIP(hipGetLastError());
    uint32_t h[2] = {0, 0};
    KC_HIP(hipMemcpyAsync(h, sm, 8, hipMemcpyDeviceToHost, s));
    KC_HIP(hipStreamSynchronize(s));
    uint64_t nn = n_in;
    if (nn == 0) nn = h[1] ? (uint64_t)h[0] + 1 : 0;
    if (h[1] && h[0] >= nn) return fail(KC_ERR_ARG, "vertex id %u >= n=%llu", h[0], (unsigned long long)nn);
    if (nn >= 0xFFFFFFFFull) return fail(KC_ERR_LIMIT, "n=%llu does not fit 32-bit ids", (unsigned long long)nn);
    const uint32_t n = (uint32_t)nn;
    DevBuf deg, off, cur, tmp, ucnt, big, scan_tmp;
    if ((rc = deg.ensure(4 * ((size_t)n + 1))) || (rc = off.ensure(4 * ((size_t)n + 1))) ||
        (rc = cur.ensure(4 * ((size_t)n + 1))) || (rc = tmp.ensure(8 * ne + 4)) ||
        (rc = ucnt.ensure(4 * ((size_t)n + 1))) || (rc = big.ensure(4 * ((size_t)n + 1))) ||
        (rc = rowptr.ensure(4 * ((size_t)n + 1))))
        return rc;
    const int gn = (int)std::max<uint64_t>(1, std::min<uint64_t>(((uint64_t)n + BUILD_T - 1) / BUILD_T, (uint64_t)ctx->cu * 16));
    hipLaunchKernelGGL(k_build_fill, dim3(gn), dim3(BUILD_T), 0, s, deg.as<uint32_t>(), (uint64_t)n + 1, 0u);
    if (ne) hipLaunchKernelGGL(k_edge_degree, dim3(ge), dim3(BUILD_T), 0, s, S, D, (uint64_t)ne, deg.as<uint32_t>());
    KC_HIP(hipGetLastError());
    if ((rc = scan_exclusive_u32(deg.as<uint32_t>(), off.as<uint32_t>(), n, scan_tmp, s))) return rc;
    KC_HIP(hipMemcpyAsync(cur.p, off.p, 4 * (size_t)n, hipMemcpyDeviceToDevice, s));
    if (ne) hipLaunchKernelGGL(k_edge_scatter, dim3(ge), dim3(BUILD_T), 0, s, S, D, (uint64_t)ne, cur.as<uint32_t>(),
                               tmp.as<uint32_t>());
    // rows: distinct sorted values at each row's start, counts in ucnt
    const int gw = (int)std::max<uint64_t>(1, std::min<uint64_t>(((uint64_t)n + 3) / 4, (uint64_t)ctx->cu * 8));
    if (n) hipLaunchKernelGGL(k_row_sort_small, dim3(gw), dim3(BUILD_T), 0, s, off.as<uint32_t>(), n,
                              tmp.as<uint32_t>(), ucnt.as<uint32_t>(), big.as<uint32_t>(), sm + 2);
    KC_HIP(hipGetLastError());
    uint32_t nbig = 0;
    KC_HIP(hipMemcpyAsync(&nbig, sm + 2, 4, hipMemcpyDeviceToHost, s));
    KC_HIP(hipStreamSynchronize(s));
    if (nbig) {
        const uint32_t words = (n + 63) / 64;
        // one bitmap slice per block, at most 1 GiB of them in all
        const uint64_t fit = std::max<uint64_t>(1, (1ull << 30) / (8ull * words));
        const uint32_t nb = (uint32_t)std::min<uint64_t>({(uint64_t)nbig, (uint64_t)ctx->cu, fit});
        DevBuf bm;
        if ((rc = bm.ensure(8ull * words * nb))) return rc;
        hipLaunchKernelGGL(k_row_sort_big, dim3(nb), dim3(BUILD_T), 0, s, off.as<uint32_t>(), big.as<uint32_t>(), nbig,
                           tmp.as<uint32_t>(), ucnt.as<uint32_t>(), bm.as<uint64_t>(), words);
        KC_HIP(hipGetLastError());
        KC_HIP(hipStreamSynchronize(s));  // bm goes out of scope
    }
    if ((rc = scan_exclusive_u32(ucnt.as<uint32_t>(), rowptr.as<uint32_t>(), n, scan_tmp, s))) return rc;
    uint32_t nnz = 0;
    KC_HIP(hipMemcpyAsync(&nnz, rowptr.as<uint32_t>() + n, 4, hipMemcpyDeviceToHost, s));
    KC_HIP(hipStreamSynchronize(s));
    if ((rc = col.ensure(4 * (size_t)nnz + 4))) return rc;
    if (n) {
        hipLaunchKernelGGL(k_row_compact, dim3(gw), dim3(BUILD_T), 0, s, off.as<uint32_t>(), tmp.as<uint32_t>(),
                           rowptr.as<uint32_t>(), n, col.as<uint32_t>());
        hipLaunchKernelGGL(k_build_max, dim3(gn), dim3(BUILD_T), 0, s, ucnt.as<uint32_t>(), n, sm + 3);
    }
    KC_HIP(hipGetLastError());
    uint32_t md = 0;
    KC_HIP(hipMemcpyAsync(&md, sm + 3, 4, hipMemcpyDeviceToHost, s));
    KC_HIP(hipStreamSynchronize(s));
    n_out = n;
    nnz_out = nnz;
    maxdeg_out = md;
    return KC_OK;
}

}  // namespace kc
