// erd_run_requests_noalloc: the batched-request entry point (erd_run_requests)
// for callers that cannot allocate device memory (code under HIP graph
// capture, tight serving loops, harnesses that forbid hipMalloc). The segment
// program that erd_run_requests uploads into a stream-ordered allocation goes
// into `scratch`, a file-scope device array of this library, instead; the
// kernels are the same, so the results are identical.
//
// Each GPU holds its own copy of `scratch` until the process ends. Its device
// address is looked up once per device, on the first call made with that
// device current, and only that address is used from host code. Calls that
// overlap on one GPU (different streams, or threads) share the array and race;
// calls on one stream are ordered. The host staging below is serialised by a
// mutex.
#include <algorithm>
#include <mutex>
#include <utility>
#include "eigenrand_amd.h"
#include "erd_capi_internal.hpp"
#include "erd_internal.hpp"
#include "erd_kernels.hpp"
#include "erd_segment.hpp"

namespace {

using erd::Segment;

constexpr int64_t kMaxRequests = 4096;  // ERD_NOALLOC_MAX_REQUESTS
constexpr size_t kSegBytes = sizeof(Segment) * (size_t)kMaxRequests;
constexpr size_t kIdxBytes = sizeof(int32_t) * (size_t)(2 * kMaxRequests + 1);
constexpr size_t kScratchBytes = (kSegBytes + kIdxBytes + 255) & ~(size_t)255;
static_assert(kMaxRequests == ERD_NOALLOC_MAX_REQUESTS, "limit in the header");

// The segment program: segments [0, n), then engine_of[groups], begin[groups + 1].
static __device__ __align__(256) char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;                              // guards everything below
void* g_scratch[kMaxDevices] = {};            // device address of `scratch`, per device
alignas(256) char g_host[kScratchBytes];      // host image of the program
std::pair<int32_t, int32_t> g_order[kMaxRequests];  // (engine, request) sort keys

// Device address of this device's `scratch` (looked up on first use).
int scratch_of_current_device(char** out) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return ERD_ERR_HIP;
    if (dev < 0 || dev >= kMaxDevices) return ERD_ERR_UNSUPPORTED;
    if (!g_scratch[dev]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return ERD_ERR_HIP;
        g_scratch[dev] = p;
    }
    *out = static_cast<char*>(g_scratch[dev]);
    return ERD_OK;
}

}  // namespace

extern "C" int erd_run_requests_noalloc(erd_engines* e, const erd_request* reqs, int64_t n, void* stream) {
    if (layout_unsupported(e)) return ERD_ERR_UNSUPPORTED;
    if (!e || !e->states || e->count <= 0 || n < 0 || (n > 0 && !reqs)) return ERD_ERR_INVALID;
    if (n > kMaxRequests) return ERD_ERR_INVALID;  // beyond the scratch array
    if (n == 0) return ERD_OK;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    // the tangent samplers' table was built with the engines; never build it here
    if (!erd::lgamma_table_built()) return ERD_ERR_UNSUPPORTED;
    std::lock_guard<std::mutex> lk(g_mu);
    char* dev = nullptr;
    int rc = scratch_of_current_device(&dev);
    if (rc) return rc;
    // Requests grouped by engine, each engine's requests in their given order
    // (sorting on (engine, index) is stable and needs no buffer).
    for (int64_t i = 0; i < n; ++i) {
        if (reqs[i].engine < 0 || reqs[i].engine >= e->count) return ERD_ERR_INVALID;
        g_order[i] = {reqs[i].engine, (int32_t)i};
    }
    std::sort(g_order, g_order + n);
    Segment* segs = reinterpret_cast<Segment*>(g_host);
    int32_t* engine_of = reinterpret_cast<int32_t*>(g_host + sizeof(Segment) * (size_t)n);
    int32_t groups = 0;
    for (int64_t k = 0; k < n; ++k) {
        rc = erd::make_segment(reqs[g_order[k].second], segs[k]);
        if (rc) return rc;
        if (groups == 0 || engine_of[groups - 1] != g_order[k].first) engine_of[groups++] = g_order[k].first;
    }
    int32_t* begin = engine_of + groups;
    int32_t g = 0;
    for (int64_t k = 0; k < n; ++k)
        if (k == 0 || g_order[k].first != g_order[k - 1].first) begin[g++] = (int32_t)k;
    begin[groups] = (int32_t)n;
    const size_t seg_bytes = sizeof(Segment) * (size_t)n;
    const size_t bytes = seg_bytes + sizeof(int32_t) * (size_t)(2 * groups + 1);
    // From pageable memory the copy is staged before the call returns, so
    // the host image may be rewritten by the next call.
    if (hipMemcpyAsync(dev, g_host, bytes, hipMemcpyHostToDevice, s) != hipSuccess) return ERD_ERR_HIP;
    const int32_t* d_engine_of = reinterpret_cast<const int32_t*>(dev + seg_bytes);
    const hipError_t err = erd::launch_program(e->states, groups, d_engine_of, d_engine_of + groups,
                                               reinterpret_cast<const Segment*>(dev), guard_of(e), s);
    return err == hipSuccess ? ERD_OK : ERD_ERR_HIP;
}
