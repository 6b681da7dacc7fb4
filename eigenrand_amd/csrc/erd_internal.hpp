// Internal host-side helpers shared by the C ABI sources (not exported).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <vector>
#include "erd_capi_internal.hpp"
#include "erd_mv.hpp"
#include "erd_segment.hpp"

struct erd_request;

namespace erd {

// Enqueues a host-to-device copy of `bytes` on `s` through pinned staging
// (staging.hip): `src` may be released as soon as this returns, and the host
// does not wait for the stream.
hipError_t upload_async(void* dst, const void* src, size_t bytes, hipStream_t s);

// Per-call device scratch (staging.hip). hipFreeAsync of a large block can
// hold up the calling thread for up to milliseconds, and the multivariate
// calls free several GB-sized blocks per call, so released scratch goes to a
// per-device cache instead: a block is tagged with an event recorded on its
// stream at release and is handed out again, without a driver call, to the
// same stream at once (stream order makes that safe) or to another stream
// once its event has completed. Blocks come in size classes 2^k * {4,5,6,7}/4
// (small ones: 4 KiB). Cached blocks stay allocated until an allocation
// fails, which releases every idle block and retries.
hipError_t scratch_alloc(void** p, size_t bytes, hipStream_t s);
void scratch_free(void* p, size_t bytes, hipStream_t s);

// A second stream of the current device (created on first use, default
// priority, not synchronising with the null stream) for work a call splits
// in two; stream_after(to, from) makes `to` wait for what `from` holds so
// far (an event recorded on `from`).
hipError_t side_stream(hipStream_t* out);
hipError_t stream_after(hipStream_t to, hipStream_t from);

// A source of device blocks that are recycled by their owner rather than
// freed through the driver (the serving loop's block pool).
struct Recycler {
    virtual void* get(size_t bytes) = 0;
    virtual void put(void* p) = 0;

protected:
    ~Recycler() = default;
};

// Stream-ordered scratch released when the object goes out of scope. With a
// recycler, blocks come from it and go back to it on destruction, which the
// owner does only once the stream's work on them has completed.
struct Scratch {
    hipStream_t s;
    std::vector<std::pair<void*, size_t>> ptrs;
    Recycler* rec = nullptr;
    explicit Scratch(hipStream_t st, Recycler* r = nullptr) : s(st), rec(r) {}
    ~Scratch() {
        for (auto& pb : ptrs) {
            if (rec) rec->put(pb.first);
            else scratch_free(pb.first, pb.second, s);
        }
    }
    template <class T>
    hipError_t alloc(T** p, size_t count) {
        *p = nullptr;
        if (count == 0) count = 1;
        hipError_t e;
        if (rec) {
            *p = (T*)rec->get(sizeof(T) * count);
            e = *p ? hipSuccess : hipErrorOutOfMemory;
        } else {
            e = scratch_alloc((void**)p, sizeof(T) * count, s);
        }
        if (e == hipSuccess) ptrs.emplace_back(*p, sizeof(T) * count);
        return e;
    }
    // Keep an allocation past this scope: the caller releases it later with
    // scratch_free (same size) or hipFreeAsync.
    void keep(const void* p) {
        for (size_t i = 0; i < ptrs.size(); ++i)
            if (ptrs[i].first == p) {
                ptrs.erase(ptrs.begin() + (long)i);
                return;
            }
    }
    // Device copy of a host vector (which may be released on return).
    template <class T>
    hipError_t upload(T** p, const std::vector<T>& v) {
        hipError_t e = alloc(p, v.size());
        if (e == hipSuccess && !v.empty()) e = upload_async(*p, v.data(), sizeof(T) * v.size(), s);
        return e;
    }
};

// Host side of DiscreteGen<int32_t, int32_t>'s constructor (Discrete.h:
// 367-391) and AliasMethod::buildTable (Discrete.h:78-192), in double.
struct HostTable {
    std::vector<uint32_t> cdf;
    std::vector<int32_t> prob, alias;
    uint32_t bitmask = 0;
    // DiscreteGen<int32_t, float> (Discrete.h:492-533): the same tables in float
    std::vector<float> fcdf, fprob;
    // DiscreteGen<int32_t, double> (Discrete.h:628-649): in double
    std::vector<double> dcdf, dprob;
};

// Segment programs (capi.hip).
// p8: engines are P8_mt19937_64 adaptors (erd_engines_create_p8); p8_mode >= 0
// forces the scalar packing (P8Feed), else each segment kind's own.
int run_segments(erd_engines* e, const std::vector<Segment>& segs, const std::vector<int32_t>& engine_of,
                 hipStream_t stream, bool p8 = false, int p8_mode = -1);
int make_segment(const erd_request& r, Segment& s);

// Multivariate helpers (mv_host.hip).
int run_lz(const std::vector<MvInstance>& inst, MvInstance* d_inst, hipStream_t s, Scratch& sc,
           bool bartlett = false);
// One persistent L*Z launch: variant, grid and its tile queue.
struct LzLaunch {
    int variant = 0;
    int grid = 0;
    std::vector<MvTile> tiles;
    std::vector<int32_t> qbeg;  // kLzQueues + 1 queue bounds in tiles
};
int plan_lz(const std::vector<MvInstance>& inst, std::vector<LzLaunch>& plan);
int factor_and_split(const std::vector<int32_t>& dims, double* d_a, const std::vector<int64_t>& off,
                     uint16_t** lhi, uint16_t** llo, float** l32, std::vector<int64_t>& split_off,
                     std::vector<int32_t>* info_out, hipStream_t s, Scratch& sc);
int build_table(const double* w, int64_t K, HostTable& t);
int upload_table(const HostTable& t, Segment& sg, Scratch& sc);
// MultinomialGen's conditional-binomial rows (Multinomial.h:74-91): the float
// probability of rows 0..K-2, from weights normalised in double.
std::vector<float> multinomial_row_p(const double* w, int64_t K);

}  // namespace erd
