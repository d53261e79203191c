// kc_device.h -- device-side declarations shared by the HIP translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
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
};

// Select and check the device (gfx950).  Returns KC_OK or KC_ERR_DEVICE.
int device_ctx(DeviceCtx **out);

// Grow-only device buffer.
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    int ensure(size_t want);
    void release();
    ~DevBuf() { release(); }
    template <class T> T *as() const { return static_cast<T *>(p); }
};

// Exclusive scan of n uint32 counts into out[0..n] (out[n] = total); total
// must fit 32 bits.  in and out may alias.  tmp is scratch owned by caller.
int scan_exclusive_u32(const uint32_t *in, uint32_t *out, uint32_t n, DevBuf &tmp, hipStream_t s);

}  // namespace kc

struct kc_oriented {
    int dev = 0;
    hipStream_t stream = nullptr;
    uint32_t n = 0;
    uint64_t m = 0;
    uint32_t max_out = 0;
    uint32_t max_core = 0;
    int mode = 0;
    double pre_ms = 0;
    kc::DevBuf rowptr, col, src, key;
    // count scratch (grow-only, reused across calls on this handle)
    kc::DevBuf s_tasks, s_meta, s_binom, s_gscr, s_scan, s_tsz;
    std::vector<uint32_t> h_rowptr_cache;  // filled lazily by host-side autotune
    ~kc_oriented();
};
