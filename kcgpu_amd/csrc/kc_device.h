// kc_device.h -- device-side declarations shared by the HIP translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <memory>
#include <vector>

#include "kc_internal.h"

#define KC_HIP(call)                                                                         \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess)                                                                \
            return kc::fail(KC_ERR_DEVICE, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                            __FILE__, __LINE__);                                             \
    } while (0)

namespace kc {

struct DeviceCtx {
    int dev = -1;
    int cu = 0;
    size_t lds_per_block = 0;
    size_t hbm = 0;
    hipDeviceProp_t prop{};
    hipStream_t stream = nullptr;  // library stream (default priority), shared by handles
    // streams (default priority) on which the size-class launches of one count
    // run side by side, so that one class's last long tasks overlap the next
    // class instead of leaving the device idle; joined back into `stream`
    static constexpr int NSIDE = 3;
    hipStream_t side[NSIDE] = {nullptr, nullptr, nullptr};
};

// Select and check the device (gfx950).  Returns KC_OK or KC_ERR_DEVICE.
int device_ctx(DeviceCtx **out);

// Grow-only device buffer, or a view into memory owned elsewhere (owned =
// false; ensure() on a view checks its size and never allocates).
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    bool owned = true;
    int ensure(size_t want);
    void release();
    void view(void *q, size_t b)
    {
        release();
        p = q;
        bytes = b;
        owned = false;
    }
    ~DevBuf() { release(); }
    template <class T> T *as() const { return static_cast<T *>(p); }
};

// Exclusive scan of n uint32 counts into out[0..n] (out[n] = total); total
// must fit 32 bits.  in and out may alias.  tmp is scratch owned by caller.
int scan_exclusive_u32(const uint32_t *in, uint32_t *out, uint32_t n, DevBuf &tmp, hipStream_t s);

// scratch of one launch sequence (grow-only), or views into a fixed region
// (fixed = true: ensure() never allocates, run_queries creates no events)
struct Scratch {
    DevBuf tasks, meta, binom, gscr, size, qdesc, local;
    bool fixed = false;
    // what `size` holds when it was filled for one query: the oriented graph
    // (its uid), the granularity and the task count; task sizes (d+(u), or
    // |N+(u) & N+(v)|) do not depend on k, so the next query on that graph
    // reuses them
    uint64_t size_uid = 0;
    uint32_t size_gran = 0;
    uint64_t size_total = 0;
};

// a number no other oriented graph of the process carries
inline uint64_t next_uid()
{
    static std::atomic<uint64_t> n{1};
    return n.fetch_add(1);
}

struct QueryRun {
    const kc_oriented *o;
    kc_query q;
    uint32_t sample = 1;  // count only the tasks whose id % sample == offset (the measured autotuner,
    uint32_t offset = 0;  // kc_count_sampled)
};

struct LaunchStats {
    double count_ms = 0, prep_ms = 0, dominant_ms = 0;
    uint64_t tasks = 0, alg_bytes = 0, tree_nodes = 0, dominant_bytes = 0;
    uint32_t dominant_class = 0;
    double class_ms[6] = {0}, class_max_task_ms[6] = {0}, class_task_ms_sum[6] = {0};
    uint64_t class_bytes[6] = {0}, class_tasks[6] = {0};
    uint32_t class_kernel[6] = {0};
};

// Count every query of qs in persistent launches on stream s (all graphs on
// the current device): one launch per (algorithm, size class) bin.  With
// local != null (one query only) the per-vertex counts are added into
// local[0..n) on the device as well.
// grid_share > 1 caps every persistent launch at 1/grid_share of the
// device's resident blocks, so that launches of that many streams run side
// by side (the server's workers) instead of the first one filling the GPU.
int run_queries(Scratch &scr, hipStream_t s, const std::vector<QueryRun> &qs, std::vector<kc_result> &res,
                LaunchStats *ls, uint64_t *local = nullptr, uint32_t grid_share = 1);

// query validation and the autotuner's choice of algo / gran / P for a query
int check_query(const kc_query *q);
kc_query tune_query(const kc_oriented *o, const kc_query *q);

// copy one query's result and the launch statistics into the ABI's kc_result
void fill_result(kc_result *r, const kc_result &res, const LaunchStats &ls);

}  // namespace kc

struct kc_oriented {
    uint64_t uid = kc::next_uid();
    int dev = 0;
    hipStream_t stream = nullptr;  // the device's library stream (not owned)
    uint32_t n = 0;
    uint64_t m = 0;
    uint32_t max_out = 0;
    uint32_t max_core = 0;
    int mode = 0;
    double pre_ms = 0;
    kc::DevBuf rowptr, col, src, key;
    std::shared_ptr<kc::DevBuf> arena;  // owner of rowptr/col/src/key when they are views (batched orientation)
    // count scratch (grow-only, reused across calls on this handle)
    kc::Scratch scr, batch_scr;
    std::vector<uint32_t> h_rowptr_cache;  // filled lazily by host-side autotune
    ~kc_oriented();
};
