// Slot view of an engine held in LDS, shared by the rejection and lockstep
// kernels.
//
// Distributions consume the engine either by full draws (operator(), a slot
// = the 4 u64 lanes of one draw) or by half draws (half(), a slot = 2 u64
// lanes: low half of a draw, then its cached high half; PacketRandomEngine.h
// :325-336). In LDS the block of 312 draws is packet-major, so block slot b
// of width SW words is simply mt[SW*b .. SW*b+SW-1] in both modes.
//
// Consumption is a virtual slot sequence: `nc` carried slots (words kept
// across a twist, or the engine's cached half at the start) followed by block
// slots [pos, SLOTS). Carried slots sit in a small LDS array and are stored
// already tempered (the engine's cached half is a tempered value); block words
// are tempered as they are read.
#pragma once
#include "erd_device.hpp"

namespace erd {

template <int SW>  // words per slot: 4 (full draw) or 2 (half draw)
struct SlotMode {
    static constexpr int kSlots = kWords / SW;  // 312 or 624 per block
};

// Tempered word p of virtual slot v (v < nc: carried; else block slot pos + v - nc).
template <int SW>
__device__ __forceinline__ uint64_t slot_word(const uint64_t* mt, const uint64_t* carry, int nc,
                                              int pos, int v, int p) {
    return (v < nc) ? carry[v * SW + p] : temper(mt[(pos + v - nc) * SW + p]);
}

// Move virtual slots [from, avail) into the carry area (at most MAXC slots)
// before a twist overwrites the block. Every thread of the wave calls it.
template <int SW, int MAXC>
__device__ __forceinline__ void carry_tail(const uint64_t* mt, uint64_t* carry, int nc, int pos,
                                          int from, int avail, int lane) {
    constexpr int N = MAXC * SW;
    constexpr int PER = (N + 63) / 64;
    const int left = avail - from;
    uint64_t keep[PER];
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const int i = lane + 64 * k;
        keep[k] = 0;
        if (i < left * SW) keep[k] = slot_word<SW>(mt, carry, nc, pos, from + i / SW, i % SW);
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const int i = lane + 64 * k;
        if (i < left * SW) carry[i] = keep[k];
    }
    __syncthreads();
}

__device__ __forceinline__ uint64_t lanemask_lt(int lane) {
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}
// acc + the number of set bits of m in lanes below the calling thread
// (v_mbcnt_lo/hi).
__device__ __forceinline__ uint32_t mbcnt64(uint64_t m, uint32_t acc) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, acc));
}

}  // namespace erd
