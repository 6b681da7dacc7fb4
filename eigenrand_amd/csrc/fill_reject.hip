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
#include "erd_segment.hpp"

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

// ------------------------------------------- gamma rows on one engine ---
// GammaGen fills (fresh generators) that run one after another on one engine
// (the category rows of a Dirichlet instance), served by a workgroup of
// kRowWaves waves: op_reject_wg for the rejection regimes, op_fixed_wg and
// op_fixed_rt_wg for alpha = 1 and integer alpha < 20. The outputs and the
// final engine state are those of the same fills on one wave (run_gamma of
// the segment program); a large-K instance's serial chain runs about twice
// as fast, and the kernel keeps only the gamma paths, so it has no spills.
constexpr int kRowWaves = 4;
constexpr int kRowThreads = 64 * kRowWaves;

union GammaRowsLds {
    RejectWgLds<GammaGS, kRowWaves> gs;
    RejectWgLds<GammaTangent, kRowWaves> gt;
    uint64_t carry_fixed[kLanes * 18];  // up to 19 draws per unit
};

__device__ void run_gamma_rows(Eng& e, GammaRowsLds& s, float* out, int64_t n, float alpha, float beta, Guard gd,
                               int tid) {
    if (alpha < 1.0f) {
        GammaGS pol{beta, (float)(kE / (alpha + kE)), 1 / alpha, alpha - 1};
        op_reject_wg<GammaGS, kRowWaves>(e, s.gs, out, n, pol, gd, tid);
        return;
    }
    if (alpha == 1.0f) {
        ExponentialF32 d{out, 1.0f, beta, 1};
        op_fixed_wg<kRowThreads>(e, s.carry_fixed, n, d, tid);
        return;
    }
    const int count = (int)alpha;
    if ((float)count == alpha && count < 20) {
        op_fixed_rt_wg<kRowThreads, 18>(e, s.carry_fixed, n, count, GammaIntRtF32{out, beta, count}, tid);
        return;
    }
    GammaTangent pol{beta, sqrtf(2 * alpha - 1), alpha - 1};
    op_reject_wg<GammaTangent, kRowWaves>(e, s.gt, out, n, pol, gd, tid);
}

// Workgroup b runs the rows of group rows[b] on its engine: row k is
// `count` gammas of shape alpha[alpha_off + k] and scale 1 at out + k * count.
__global__ __launch_bounds__(kRowThreads) void k_gamma_rows(EngineState* __restrict__ states,
                                                            const GammaRowGroup* __restrict__ rows,
                                                            const float* __restrict__ alpha, Guard gd) {
    __shared__ uint64_t mt[kWords];
    __shared__ GammaRowsLds s;
    const int tid = threadIdx.x;
    const GammaRowGroup gr = rows[blockIdx.x];
    const int wl = tid < 64 ? tid : (1 << 20);  // state copies by wave 0
    Eng e;
    eng_load(e, mt, states + gr.engine, wl);
    for (int k = 0; k < gr.rows; ++k)
        run_gamma_rows(e, s, gr.out + (int64_t)k * gr.count, gr.count, alpha[gr.alpha_off + k], gr.beta, gd, tid);
    eng_store(e, states + gr.engine, wl);
}

hipError_t launch_gamma_rows(EngineState* st, const GammaRowGroup* rows, int32_t n_groups, const float* alpha,
                             Guard gd, hipStream_t s) {
    if (n_groups <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_gamma_rows, dim3((unsigned)n_groups), dim3(kRowThreads), 0, s, st, rows, alpha, gd);
    return hipGetLastError();
}

}  // namespace erd
