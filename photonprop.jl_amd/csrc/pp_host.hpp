// pp_host.hpp — host-side infrastructure shared by the library's translation
// units: error reporting, device buffers, the request layout the kernels
// read, and the context object behind pp_ctx*.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "pp_types.hpp"

namespace pph {

#define PP_HIP(call)                                                                      \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess)                                                             \
            return ::pph::fail(e_ == hipErrorOutOfMemory ? PP_ERR_OOM : PP_ERR_DEVICE,    \
                               "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_),     \
                               __FILE__, __LINE__);                                       \
    } while (0)

// Grow-only device allocation.
struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
};

inline int ensure(DevBuf& b, size_t bytes) {
    if (bytes <= b.cap) return PP_OK;
    if (b.p) PP_HIP(hipFree(b.p));
    b.p = nullptr;
    b.cap = 0;
    size_t want = std::max<size_t>(bytes, 256);
    PP_HIP(hipMalloc(&b.p, want));
    b.cap = want;
    return PP_OK;
}

}  // namespace pph

struct pp_ctx {
    int device = 0;
    int n_cu = 0;
    int blocks_per_cu = 1;
    char name[256] = {0};
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    float last_ms = 0.0f;
    bool timed = false;
    // medium
    pph::DevBuf table;
    uint32_t n_points = 0;
    float lmin = 0.0f, lmax = 0.0f, inv_dl = 0.0f, g = 0.0f, n_min = 0.0f;
    pph::DevBuf knots;
    // work buffers
    pph::DevBuf reqs, tile_start, partials, counters, staging, cnt, fill, sums, offsets, out_hits;
    pph::DevBuf rec_req;  // request of each staged record, written by the count pass
    pph::DevBuf scratch;
    // pinned staging for hit records going to the caller's host memory
    void* pinned[2] = {nullptr, nullptr};
    // pinned buffer for a single request's upload and its summary download
    // (pp_propagate), so that neither goes through the runtime's pageable path
    void* io = nullptr;
    bool occ_known = false;  // blocks_per_cu queried
    hipEvent_t copy_ev[2] = {nullptr, nullptr};
    std::vector<pph::DevReq> h_reqs;
    std::vector<uint64_t> h_tiles;
    // device address of this device's `scratch` array (allocation-free entry point)
    void* na_scratch = nullptr;
};
