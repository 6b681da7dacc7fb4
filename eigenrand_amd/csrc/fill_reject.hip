// Rejection samplers that queue accepted lanes through the reference's
// CompressMask (arch/AVX/PacketFilter.h:147-184 for 8 float lanes,
// arch/SSE/PacketFilter.h:118-150 for 4): GammaGen with alpha < 1 (GS) or
// non-integer alpha > 1 (tangent), PoissonGen with mean >= 12 and BinomialGen
// in its tangent regime.
//
// Every candidate packet consumes a fixed number of engine slots, so
// candidates are evaluated in parallel: one wave owns one engine and takes 16
// (W=8) or 32 (W=4) candidates per step, each thread evaluating two lanes of
// one candidate. The queue order is then reproduced with wave ballots:
//   - a candidate whose W lanes are all accepted is emitted as a packet of
//     its own (it bypasses the queue),
//   - accepted lanes of the other candidates form the queue in (candidate,
//     lane) order; queue item f belongs to queued packet f / W,
//   - output packet index = (#bypassing candidates before the emitting
//     candidate) + (#queued packets before it).
// mbcnt over the ballots gives each accepted lane its queue index; a small
// LDS table maps queued packets completed in this step to output packets;
// lanes whose packet is not complete yet wait in the LDS rest queue. The fill
// stops at the candidate that emits the last packet, and the engine is left
// exactly after that candidate's slots, as the reference leaves it.
#include <mutex>
#include "erd_device.hpp"
#include "erd_dists.hpp"
#include "erd_kernels.hpp"
#include "erd_ops.hpp"

namespace erd {

// ----------------------------------------------------- plgamma table ---
__global__ __launch_bounds__(256) void k_lgamma_table(float* __restrict__ lg) {
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k < kLgTab) lg[k] = plgamma_approx((float)(k + 1));
}

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_lg_mutex;
const float* g_lg[kMaxDevices] = {};
}  // namespace

const float* lgamma_table(hipStream_t s) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_lg_mutex);
    if (!g_lg[dev]) {
        float* lg = nullptr;
        if (hipMalloc(&lg, sizeof(float) * kLgTab) != hipSuccess) return nullptr;
        hipLaunchKernelGGL(k_lgamma_table, dim3(kLgTab / 256), dim3(256), 0, s, lg);
        // the table is complete before any stream of this device can use it
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s) != hipSuccess) {
            (void)hipFree(lg);
            return nullptr;
        }
        g_lg[dev] = lg;
    }
    return g_lg[dev];
}

const float* lgamma_table_built() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_lg_mutex);
    return g_lg[dev];
}

// gs (optional): per-engine queues of a reused generator object.
template <class P>
__global__ __launch_bounds__(64) void k_fill_reject(EngineState* __restrict__ states,
                                                    typename P::Out* __restrict__ out, int64_t per,
                                                    P pol, Guard gd,
                                                    GenState* __restrict__ gs) {
    __shared__ uint64_t mt[kWords];
    __shared__ RejectLds<P> s;
    const int lane = threadIdx.x;
    const int64_t g = blockIdx.x;
    if (per / P::W == 0) return;
    Eng e;
    eng_load(e, mt, states + g, lane);
    op_reject(e, s, out + g * per, per, pol, gd, lane, gs ? gs + g : nullptr);
    eng_store(e, states + g, lane);
}

// ------------------------------------------------------------- launchers ---
template <class P>
static hipError_t launch_reject(EngineState* st, int64_t n, typename P::Out* out, int64_t per,
                                const P& pol, Guard gd, GenState* gs, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_reject<P>, dim3((unsigned)n), dim3(64), 0, s, st, out, per, pol, gd, gs);
    return hipGetLastError();
}

hipError_t launch_fill_gamma_reject_f32(EngineState* st, int64_t n, float* out, int64_t per,
                                        float alpha, float beta, Guard gd, hipStream_t s,
                                        GenState* gs) {
    if (alpha < 1.0f) {
        GammaGS pol{beta, (float)(kE / (alpha + kE)), 1 / alpha, alpha - 1};
        return launch_reject(st, n, out, per, pol, gd, gs, s);
    }
    GammaTangent pol{beta, sqrtf(2 * alpha - 1), alpha - 1};
    return launch_reject(st, n, out, per, pol, gd, gs, s);
}

hipError_t launch_fill_poisson_reject_i32(EngineState* st, int64_t n, int32_t* out, int64_t per,
                                          const PoissonParams& pp, Guard gd, hipStream_t s,
                                          GenState* gs) {
    const float* lg = lgamma_table(s);
    if (!lg) return hipErrorOutOfMemory;
    PoissonTangent pol{(float)pp.sqrt_tmean, (float)pp.mean, (float)pp.log_mean, (float)pp.g1, lg};
    return launch_reject(st, n, out, per, pol, gd, gs, s);
}

hipError_t launch_fill_binomial_reject_i32(EngineState* st, int64_t n, int32_t* out, int64_t per,
                                           const BinomialParams& bp, Guard gd, hipStream_t s,
                                           GenState* gs) {
    const float* lg = lgamma_table(s);
    if (!lg) return hipErrorOutOfMemory;
    BinomialTangent pol{(float)bp.trials, (float)bp.sqrt_v, (float)bp.mean, (float)bp.log_small_p,
                        (float)bp.log_small_q, (float)bp.g1, bp.trials, bp.p == bp.small_p ? 0 : 1, lg};
    return launch_reject(st, n, out, per, pol, gd, gs, s);
}

}  // namespace erd
