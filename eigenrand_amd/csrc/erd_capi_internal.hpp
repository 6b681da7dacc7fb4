// Definition of the opaque erd_engines handle (shared by the C ABI sources).
#pragma once
#include <stdint.h>
#include <vector>
#include "erd_device.hpp"

struct erd_engines {
    erd::EngineState* states = nullptr;  // device records
    uint32_t* status = nullptr;          // device status word (kernel error bits)
    int64_t count = 0;
    bool owns = false;                   // views share the parent's memory
    int32_t reject_cap = erd::kDefaultRejectCap;      // per-launch iteration caps (erd::Guard)
    int32_t lockstep_cap = erd::kDefaultLockstepCap;
    int32_t layout = 8;  // float packet width of the fills: 4 (SSE2/NEON), 8 (AVX2) or 16 (AVX-512)
};

// Calls that exist only at the AVX2 layout refuse a set with another one.
inline bool layout_unsupported(const erd_engines* e) { return e && e->layout != 8; }

// Status word and iteration caps of the launches on an engine set.
inline erd::Guard guard_of(const erd_engines* e) { return {e->status, e->reject_cap, e->lockstep_cap}; }

// Factors of a batch of MvNormalGen / WishartGen instances: the FP64 Cholesky
// of each covariance, kept as bf16 hi/lo planes (row-major, padded to a
// multiple of 16) for the L*Z transform.
struct erd_mv_factors {
    std::vector<int32_t> dims;
    std::vector<int64_t> split_off;  // plane offset of instance i (elements)
    uint16_t* lhi = nullptr;         // device, owned
    uint16_t* llo = nullptr;
};
