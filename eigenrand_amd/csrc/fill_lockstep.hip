// Packet-lockstep samplers on half draws (Packet4f through half(),
// PacketRandomEngine.h:325-336): a packet of 4 int32 consumes a data-dependent
// number of slots, so packets are produced one after another per engine.
//
//  - PoissonGen, mean < 12 (Discrete.h:860-873): Knuth's product loop runs
//    until no lane's product exceeds e^-mean; the packet consumes
//    1 + max(lane counts) half draws. Also BinomialGen with n >= 25 and
//    n*min(p,1-p) < 1 (Discrete.h:1007-1010), plus its final flip.
//  - BinomialGen, n < 25 (Discrete.h:996-1006): n half draws per packet, each
//    lane counts draws >= p. Slots are independent, so up to 16 are compared
//    at once by the whole wave and counted with ballots.
//  - BinomialVGen<int32_t> (Discrete.h:1124-1192): per-lane parameters; the
//    lanes with mean < 1 run the Knuth loop in lockstep, then the others run
//    tangent rejection keeping each lane's first accepted candidate. Sixteen
//    candidate iterations are evaluated at once (thread 4i+j runs iteration i
//    of lane j); a ballot per lane finds its first acceptance, and the packet
//    consumes exactly the iterations up to the slowest lane, as the sequential
//    loop would.
#include <cmath>
#include "erd_device.hpp"
#include "erd_kernels.hpp"
#include "erd_ops.hpp"

namespace erd {

// PoissonGen (mean < 12) and BinomialGen's Poisson regime (flip applied).
__global__ __launch_bounds__(64) void k_poisson_knuth(EngineState* __restrict__ states,
                                                      int32_t* __restrict__ out, int64_t per,
                                                      float thr, int32_t trials, int flip,
                                                      Guard gd) {
    __shared__ uint64_t mt[kWords];
    __shared__ uint64_t carry[2];
    __shared__ __align__(16) float ustage[kKnuthStage];
    const int lane = threadIdx.x;
    const int64_t g = blockIdx.x;
    if (per / 4 == 0) return;
    Eng e;
    eng_load(e, mt, states + g, lane);
    op_knuth(e, carry, ustage, out + g * per, per, thr, trials, flip, gd, lane);
    eng_store(e, states + g, lane);
}

// The same for thresholds below kKnuthPrefixThr: the prefix path keeps its
// window in registers, so the workgroup needs only the engine in LDS.
__global__ __launch_bounds__(64) void k_poisson_knuth_lp(EngineState* __restrict__ states,
                                                         int32_t* __restrict__ out, int64_t per,
                                                         float thr, int32_t trials, int flip,
                                                         Guard gd) {
    __shared__ uint64_t mt[kWords];
    __shared__ uint64_t carry[2];
    const int lane = threadIdx.x;
    const int64_t g = blockIdx.x;
    if (per / 4 == 0) return;
    Eng e;
    eng_load(e, mt, states + g, lane);
    op_knuth_lp(e, carry, out + g * per, per / 4, thr, trials, flip, gd, lane);
    eng_store(e, states + g, lane);
}

// BinomialGen, trials < 25.
__global__ __launch_bounds__(64) void k_binomial_small(EngineState* __restrict__ states,
                                                       int32_t* __restrict__ out, int64_t per,
                                                       int32_t trials, float pp) {
    __shared__ uint64_t mt[kWords];
    __shared__ uint64_t carry[2 * kBinSmallMaxCarry];
    const int lane = threadIdx.x;
    const int64_t g = blockIdx.x;
    if (per / 4 == 0) return;
    Eng e;
    eng_load(e, mt, states + g, lane);
    op_binomial_small(e, carry, out + g * per, per, trials, pp, lane);
    eng_store(e, states + g, lane);
}

// BinomialVGen<int32_t>, per-lane float parameters.
__global__ __launch_bounds__(64) void k_binomial_vop(EngineState* __restrict__ states,
                                                     int32_t* __restrict__ out, int64_t per,
                                                     Param<int32_t> trials, Param<float> prob,
                                                     Guard gd,
                                                     const float* __restrict__ lg) {
    __shared__ uint64_t mt[kWords];
    __shared__ uint64_t carry[2];
    const int lane = threadIdx.x;
    const int64_t g = blockIdx.x;
    if (per / 4 == 0) return;
    Eng e;
    eng_load(e, mt, states + g, lane);
    op_binomial_vop(e, carry, out + g * per, per, trials + g * per, prob + g * per, gd, lg,
                    lane);
    eng_store(e, states + g, lane);
}

// BinomialVGen<float>: Packet8f lanes on full draws, per-element trials and p.
__global__ __launch_bounds__(64) void k_binomial_vopf(EngineState* __restrict__ states,
                                                      int32_t* __restrict__ out, int64_t per,
                                                      Param<int32_t> trials, Param<float> prob,
                                                      Guard gd,
                                                      const float* __restrict__ lg) {
    __shared__ uint64_t mt[kWords];
    __shared__ uint64_t carry[2 * kLanes];
    const int lane = threadIdx.x;
    const int64_t g = blockIdx.x;
    const int64_t T = per / 8;
    if (T == 0) return;
    Eng e;
    eng_load(e, mt, states + g, lane);
    BinomialVopFSrc src{out + g * per, trials + g * per, prob + g * per, per, 0, 0.f, 0.f, 0.f, 0.f};
    op_binomial_v8(e, carry, T, src, gd, lg, lane);
    eng_store(e, states + g, lane);
}

// Vectorize-over-params gamma (Packet8f) and Poisson (Packet4f).
__global__ __launch_bounds__(64) void k_gamma_vop(EngineState* __restrict__ states,
                                                  float* __restrict__ out, int64_t per,
                                                  const float* __restrict__ alpha,
                                                  const float* __restrict__ beta,
                                                  Guard gd) {
    __shared__ uint64_t mt[kWords];
    __shared__ uint64_t carry[2 * kLanes];
    const int lane = threadIdx.x;
    const int64_t g = blockIdx.x;
    if (per / 8 == 0) return;
    Eng e;
    eng_load(e, mt, states + g, lane);
    op_gamma_vop(e, carry, out + g * per, per, alpha + g * per, beta + g * per, gd, lane);
    eng_store(e, states + g, lane);
}

__global__ __launch_bounds__(64) void k_poisson_vop(EngineState* __restrict__ states,
                                                    int32_t* __restrict__ out, int64_t per,
                                                    const float* __restrict__ mean,
                                                    Guard gd,
                                                    const float* __restrict__ lg) {
    __shared__ uint64_t mt[kWords];
    __shared__ uint64_t carry[2];
    const int lane = threadIdx.x;
    const int64_t g = blockIdx.x;
    if (per / 4 == 0) return;
    Eng e;
    eng_load(e, mt, states + g, lane);
    op_poisson_vop(e, carry, out + g * per, per, mean + g * per, gd, lg, lane);
    eng_store(e, states + g, lane);
}

hipError_t launch_fill_gamma_vop_f32(EngineState* st, int64_t n, float* out, int64_t per,
                                     const float* alpha, const float* beta, Guard gd,
                                     hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_gamma_vop, dim3((unsigned)n), dim3(64), 0, s, st, out, per, alpha, beta,
                       gd);
    return hipGetLastError();
}

hipError_t launch_fill_poisson_vop_i32(EngineState* st, int64_t n, int32_t* out, int64_t per,
                                       const float* mean, Guard gd, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const float* lg = lgamma_table(s);
    if (!lg) return hipErrorOutOfMemory;
    hipLaunchKernelGGL(k_poisson_vop, dim3((unsigned)n), dim3(64), 0, s, st, out, per, mean, gd, lg);
    return hipGetLastError();
}

hipError_t launch_fill_poisson_knuth_i32(EngineState* st, int64_t n, int32_t* out, int64_t per,
                                         double mean, int32_t trials, int flip, Guard gd,
                                         hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const float thr = (float)std::exp(-mean);
    hipLaunchKernelGGL(thr < kKnuthPrefixThr ? k_poisson_knuth_lp : k_poisson_knuth, dim3((unsigned)n), dim3(64), 0,
                       s, st, out, per, thr, trials, flip, gd);
    return hipGetLastError();
}

hipError_t launch_fill_binomial_small_i32(EngineState* st, int64_t n, int32_t* out, int64_t per,
                                          int32_t trials, double p, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_binomial_small, dim3((unsigned)n), dim3(64), 0, s, st, out, per, trials,
                       (float)p);
    return hipGetLastError();
}

hipError_t launch_fill_binomial_vop_i32(EngineState* st, int64_t n, int32_t* out, int64_t per,
                                        Param<int32_t> trials, Param<float> p, Guard gd,
                                        hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const float* lg = lgamma_table(s);
    if (!lg) return hipErrorOutOfMemory;
    hipLaunchKernelGGL(k_binomial_vop, dim3((unsigned)n), dim3(64), 0, s, st, out, per, trials, p,
                       gd, lg);
    return hipGetLastError();
}

hipError_t launch_fill_binomial_vop_f32(EngineState* st, int64_t n, int32_t* out, int64_t per,
                                        Param<int32_t> trials, Param<float> p, Guard gd,
                                        hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const float* lg = lgamma_table(s);
    if (!lg) return hipErrorOutOfMemory;
    hipLaunchKernelGGL(k_binomial_vopf, dim3((unsigned)n), dim3(64), 0, s, st, out, per, trials, p,
                       gd, lg);
    return hipGetLastError();
}

}  // namespace erd
