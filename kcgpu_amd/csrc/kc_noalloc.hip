// kc_noalloc.hip -- kc_count_noalloc: kc_count without any device allocation.
//
// The counting launches need scratch beyond the oriented CSR and the result:
// per-task sizes (4 bytes a task), the binned task list (8 bytes a task),
// query descriptors, queue heads and accumulators, the binomial table and,
// for tasks above 1024 vertices, per-workgroup HBM bitmap slices.  kc_count
// keeps that scratch in grow-only hipMalloc buffers of the oriented handle.
// This entry point carves it out of one file-scope device array instead, so
// that a caller that cannot allocate device memory (a serving loop with a
// fixed footprint, a harness that forbids hipMalloc) can still count.
//
// The array lives in every device the library's code object is loaded on,
// for the life of the process.  Host code never names it as an address: its
// device address is looked up once per device with hipGetSymbolAddress and
// only that pointer is handed to kernels.  Calls on one device are serialized
// by a lock (they would otherwise race on the one array); the array keeps the
// last call's contents, and everything a count needs zeroed is cleared by a
// fill kernel at the start of the call.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>

#include "kc_device.h"

namespace {

constexpr size_t NOALLOC_BYTES = (size_t)KC_NOALLOC_SCRATCH_MIB << 20;
constexpr int MAX_DEVICES = 64;

static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

std::mutex g_noalloc_mu[MAX_DEVICES];
void *g_scratch_ptr[MAX_DEVICES] = {};

inline size_t up256(size_t x) { return (x + 255) & ~(size_t)255; }

// device address of `scratch` on the current device `dev`, looked up once
int scratch_ptr(int dev, unsigned char **out)
{
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return kc::fail(KC_ERR_DEVICE, "hipGetSymbolAddress(scratch) on device %d failed: %s", dev,
                            hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    *out = static_cast<unsigned char *>(g_scratch_ptr[dev]);
    return KC_OK;
}

}  // namespace

extern "C" int kc_count_noalloc(const kc_oriented *o, const kc_query *q, kc_result *r)
{
    using namespace kc;
    if (!o || !q || !r) return fail(KC_ERR_ARG, "kc_count_noalloc: null argument");
    int rc = check_query(q);
    if (rc) return rc;
    if (o->dev < 0 || o->dev >= MAX_DEVICES) return fail(KC_ERR_ARG, "kc_count_noalloc: device %d", o->dev);
    // k <= 2 are closed forms on the host (kc_count does no device work for them either)
    const kc_query qq = tune_query(o, q);
    if (qq.k <= 2) return kc_count(o, q, r);

    std::lock_guard<std::mutex> lk(g_noalloc_mu[o->dev]);
    KC_HIP(hipSetDevice(o->dev));
    unsigned char *base = nullptr;
    if ((rc = scratch_ptr(o->dev, &base))) return rc;

    // regions of the array, each 256-byte aligned, in the order run_queries
    // fills them: descriptors, counters, task sizes, task list, binomials,
    // and whatever is left for the HBM bitmap slices
    const uint64_t total = qq.gran == KC_GRAN_EDGE ? o->m : o->n;
    const size_t need_qdesc = 4096, need_meta = 64 * 1024;
    const size_t need_size = up256(4 * (size_t)(total + 1));
    const size_t need_tasks = up256(8 * (size_t)(total + 1));
    // task sizes never exceed the largest out-degree; the table holds
    // C(n, r) for n <= size + 1, r <= k, 9 bytes an entry
    const size_t need_binom = up256(((size_t)o->max_out + 2) * (qq.k + 1) * 9 + 16);
    const size_t fixed_part = need_qdesc + need_meta + need_size + need_tasks + need_binom;
    if (fixed_part > NOALLOC_BYTES)
        return fail(KC_ERR_LIMIT,
                    "kc_count_noalloc: %llu %s tasks need %zu bytes of scratch, the fixed array holds %zu "
                    "(%d MiB)",
                    (unsigned long long)total, qq.gran == KC_GRAN_EDGE ? "edge" : "node", fixed_part,
                    NOALLOC_BYTES, KC_NOALLOC_SCRATCH_MIB);
    Scratch scr;
    scr.fixed = true;
    size_t at = 0;
    auto carve = [&](DevBuf &b, size_t bytes) {
        b.view(base + at, bytes);
        at += bytes;
    };
    carve(scr.qdesc, need_qdesc);
    carve(scr.meta, need_meta);
    carve(scr.size, need_size);
    carve(scr.tasks, need_tasks);
    carve(scr.binom, need_binom);
    carve(scr.gscr, NOALLOC_BYTES - at);

    std::memset(r, 0, sizeof *r);
    std::vector<QueryRun> qs{QueryRun{o, qq}};
    std::vector<kc_result> res;
    LaunchStats ls;
    rc = run_queries(scr, o->stream, qs, res, &ls);
    if (rc) return rc;
    fill_result(r, res[0], ls);
    return r->overflow ? fail(KC_ERR_OVERFLOW, "k-clique count exceeds 2^64-1") : KC_OK;
}
