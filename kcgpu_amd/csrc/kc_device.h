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

// The simple undirected CSR of a raw edge list built on the device
// (kc_build.hip): rowptr[n+1] and col[nnz] in HBM, the same graph as the host
// builder.  n_in = 0 derives max id + 1.
int build_csr_device(DeviceCtx *ctx, hipStream_t s, uint32_t n_in, size_t ne, const uint32_t *src,
                     const uint32_t *dst, DevBuf &rowptr, DevBuf &col, uint32_t &n, uint64_t &nnz, uint32_t &maxdeg);

// Orient a simple undirected CSR already in HBM (kc_prep.hip).
int orient_device_csr(DeviceCtx *ctx, uint32_t n, uint32_t nnz, uint32_t maxdeg, const uint32_t *frow,
                      const uint32_t *fcol, int mode, kc_oriented **out);

// query validation and the autotuner's choice of algo / gran / P for a query
int check_query(const kc_query *q);
kc_query tune_query(const kc_oriented *o, const kc_query *q);

// ---- the serving loop's resident launch (open queues) ----------------------
// mailbox of a query slot, in host memory: the device writes count and ovf,
// then seq (the ticket) with release; the host waits for seq
struct OpenMail {
    uint64_t count;
    uint32_t ovf;
    uint32_t seq;
};
// descriptor of a query slot, in host memory, written before the query's
// segment is published
struct OpenSlot {
    const uint32_t *rowptr, *col, *src;
    uint32_t k, gran;
    uint64_t total;  // tasks (vertices or oriented edges)
    uint32_t seq, pad;
};
static_assert(sizeof(OpenSlot) == 48, "OpenSlot is read as six 64-bit words");
constexpr uint32_t OPEN_MAX_ROWS = 64;  // largest task the resident launches take (one-word rows)

struct OpenQueryIn {
    const kc_oriented *o;
    kc_query q;      // pivoting, k >= 3; every task of o at most OPEN_MAX_ROWS vertices
    uint32_t seq;    // ticket, non-zero
};

// Resident launches of the pivoting kernel for tasks of up to OPEN_MAX_ROWS
// vertices, each draining one queue of query segments that the host
// publishes in pinned memory while it runs: queue 0 (express) for queries
// with few tasks, on a few waves per CU, and queue 1 (bulk) on the rest.
// Not thread-safe: one host thread (the server's engine) drives it.
struct OpenEngine {
    static constexpr uint32_t SLOTS = 4096;
    static constexpr uint64_t NSEG = 1u << 16;  // segments per queue between restarts
    static constexpr int NQ = 2;
    static constexpr int CTL_WORDS = 32;  // device words per queue (counter, known, last poll): own 256 bytes
    struct Queue {
        hipStream_t run = nullptr;
        hipEvent_t done = nullptr;   // recorded after each launch
        bool launched = false, stopped = false;
        uint64_t total = 0, nseg = 0;  // task indices and segments published
        uint64_t *htotal = nullptr, *hnseg = nullptr, *hseg = nullptr;  // host (pinned)
        uint32_t blocks = 0;
        uint64_t inflight = 0;
    };
    int dev = -1;
    hipStream_t adm = nullptr;       // set-up (init) only
    Queue qs[NQ];
    DevBuf qcount, qovf, qdone, ctl, binom, acc;
    uint8_t *pinned = nullptr;       // host: queues, slots, mailboxes
    uint8_t *pinned_dev = nullptr;   // the same memory's device address
    OpenSlot *hslot = nullptr;
    OpenMail *mail = nullptr;
    size_t lds = 0;
    std::vector<uint32_t> free_slots;
    ~OpenEngine();
    int init(int dev);
    // true while queue j's launch has not ended
    bool running(int j);
    // queue j has no launch and nothing in flight: restart its indices at 0
    // (never waits on the device)
    int reset(int j);
    // publish the queries as segments of queue j; slot[i] gets query i's slot
    // (~0u when it has no task: its count is 0).  The caller checks room()
    // and free_slots first.
    void admit(int j, const std::vector<OpenQueryIn> &in, std::vector<uint32_t> &slot);
    void release(uint32_t slot) { free_slots.push_back(slot); }
    // start queue j's launch (call when it is not running)
    int launch(int j);
    // end queue j's launch: one stop index per wave
    void stop(int j);
    uint64_t room(int j) const { return NSEG - 1 - qs[j].nseg; }  // (one segment kept for the stop)
};

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
