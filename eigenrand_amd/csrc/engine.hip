// Engine-set kernels: seeding and discard.
#include "erd_device.hpp"
#include "erd_kernels.hpp"

namespace erd {

// MersenneTwister::seed (PacketRandomEngine.h:256-265): lane l of engine g is
// seeded seeds[g] + l; x[i] = f*(x[i-1] ^ x[i-1]>>62) + i. The recurrence is
// serial per lane, so four threads run it into LDS and the wave stores the
// record with 16-byte writes.
__global__ __launch_bounds__(64) void k_seed(EngineState* __restrict__ states,
                                             const uint64_t* __restrict__ seeds) {
    __shared__ uint64_t mt[kWords];
    const int lane = threadIdx.x;
    EngineState* st = states + blockIdx.x;
    if (lane < kLanes) {
        uint64_t prev = seeds[blockIdx.x] + (uint64_t)lane;
        mt[lane] = prev;
        for (int i = 1; i < kN; ++i) {
            prev = (prev ^ (prev >> 62)) * kF + (uint64_t)i;
            mt[i * kLanes + lane] = prev;
        }
    }
    __syncthreads();
    store_state(st, mt, lane);
    if (lane == 0) {
        st->idx = kN;
        st->half_valid = 0;
        st->half_cache[0] = 0;
        st->half_cache[1] = 0;
        st->reserved = 0;
    }
}

hipError_t launch_seed(EngineState* st, const uint64_t* d_seeds, int64_t n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_seed, dim3((unsigned)n), dim3(64), 0, s, st, d_seeds);
    return hipGetLastError();
}

// discard(n) (PacketRandomEngine.h:317-323): advance every engine by `calls`
// draws. Only the block position and the twists matter, so no word is tempered.
__global__ __launch_bounds__(64) void k_discard(EngineState* __restrict__ states, int64_t calls) {
    __shared__ uint64_t mt[kWords];
    const int lane = threadIdx.x;
    EngineState* st = states + blockIdx.x;
    int64_t pos = st->idx;
    const int64_t end = pos + calls;
    if (end <= kN) {
        if (lane == 0) st->idx = (int32_t)end;
        return;
    }
    load_state(mt, st, lane);
    // Draws at positions [pos, end) of the block sequence; a twist is due each
    // time the position reaches a multiple of kN.
    // (pos == kN, a spent block, is covered: its first draw is position kN.)
    const int64_t twists = (end - 1) / kN;  // blocks entered after the current one
    for (int64_t t = 0; t < twists; ++t) twist_lds(mt, lane);
    store_state(st, mt, lane);
    if (lane == 0) st->idx = (int32_t)(end - twists * kN);
}

hipError_t launch_discard(EngineState* st, int64_t n, int64_t calls, hipStream_t s) {
    if (n <= 0 || calls <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_discard, dim3((unsigned)n), dim3(64), 0, s, st, calls);
    return hipGetLastError();
}

}  // namespace erd
