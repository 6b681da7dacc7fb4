// Host-side launchers for the device kernels (one declaration per kernel
// family). All take device pointers and enqueue on `s`.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace erd {

struct EngineState;
struct Segment;
struct GenState;
struct Guard;
template <class T> struct Param;

// PoissonGen constructor constants, computed on the host in double with libm
// (Discrete.h:807-813).
struct PoissonParams {
    double mean, ne_mean, sqrt_tmean, log_mean, g1;
};
// BinomialGen constructor constants (Discrete.h:929-941).
struct BinomialParams {
    int32_t trials;
    double p, small_p, mean, g1, sqrt_v, log_small_p, log_small_q;
};

// engine.hip
hipError_t launch_seed(EngineState* st, const uint64_t* d_seeds, int64_t n, hipStream_t s);
hipError_t launch_discard(EngineState* st, int64_t n, int64_t calls, hipStream_t s);

// fill_fixed.hip
// One engine's fill of `count` floats into `out` (count a multiple of 8).
struct FillJob {
    int32_t engine;
    int32_t pad;
    int64_t count;
    float* out;
};
// Threads per job: 256, or 640 (`wide`, one item of a draw block per
// thread) when the jobs are few enough to run in one round.
constexpr int kJobThreads = 256;
constexpr int kJobThreadsWide = 640;
hipError_t launch_std_normal_jobs(EngineState* st, const FillJob* jobs, int n_jobs, hipStream_t s, bool wide);
// Consecutive GammaGen fills (fresh generators) on one engine: `rows` rows
// of `count` gammas, row k of shape alpha[alpha_off + k] and scale `beta` at
// out + k * count (a Dirichlet instance's category rows, or one fill).
struct GammaRowGroup {
    int32_t engine;
    int32_t rows;
    int64_t count;
    float* out;
    int64_t alpha_off;
    float beta;
    int32_t pad;
};
// One workgroup of four waves per group.
hipError_t launch_gamma_rows(EngineState* st, const GammaRowGroup* rows, int32_t n_groups, const float* alpha,
                             Guard gd, hipStream_t s);
hipError_t launch_fill_u64(EngineState* st, int64_t n, uint64_t* out, int64_t per, hipStream_t s);
// Fills at the SSE2/NEON (W = 4) and AVX-512 (W = 16) lane layouts whose
// element order differs from the AVX2 one. `mode`: uniform 0 = [0,1),
// 1 = a + u (b - a); normal 0 standard, 1 z*b + a (mean a, sd b), 3 lognormal.
enum LayoutKind {
    LK_U64 = 0,
    LK_UNIFORM_F32,
    LK_UNIFORM_F64,
    LK_NORMAL_F32,
    LK_EXPONENTIAL_F32,
    LK_CAUCHY_F32,
    LK_WEIBULL_F32,
    LK_EXTREME_VALUE_F32,
    LK_STUDENT_T_F32,
    LK_NORMAL_F64,
    LK_EXPONENTIAL_F64,
    LK_CAUCHY_F64
};
hipError_t launch_fill_layout(EngineState* st, int64_t n, int W, int kind, void* out, int64_t per, double a,
                              double b, int mode, hipStream_t s);
// P8 adaptor fills: n Vmt engines (n/2 adaptors), per_v outputs per Vmt
// engine; mode 0 u64, 1 u32, 2 uniform_real() floats.
hipError_t launch_fill_p8(EngineState* st, int64_t n, void* out, int64_t per_v, int mode,
                          hipStream_t s);
hipError_t launch_fill_uniform_f32(EngineState* st, int64_t n, float* out, int64_t per, float a,
                                   float b, int affine, hipStream_t s);
hipError_t launch_fill_uniform_f64(EngineState* st, int64_t n, double* out, int64_t per, double a,
                                   double b, int affine, hipStream_t s);
hipError_t launch_fill_normal_f32(EngineState* st, int64_t n, float* out, int64_t per, float mean,
                                  float sd, int affine, hipStream_t s);
// The fast precision mode of normal<float> (hardware log2/sqrt/sin/cos).
hipError_t launch_fill_normal_f32_fast(EngineState* st, int64_t n, float* out, int64_t per, float mean,
                                       float sd, int affine, hipStream_t s);
// Write-only stores of n_blocks contiguous blocks of `per` floats.
hipError_t launch_store_blocks(float* out, int64_t n_blocks, int64_t per, hipStream_t s);
// DiscreteGen<int32_t, int32_t> fill: CDF (ncdf = K < 16) or alias table.
hipError_t launch_fill_discrete_i32(EngineState* st, int64_t n, int32_t* out, int64_t per, const uint32_t* cdf,
                                   const int32_t* prob, const int32_t* alias, int32_t ncdf, uint32_t bitmask,
                                   hipStream_t s);
hipError_t launch_fill_normal_f64(EngineState* st, int64_t n, double* out, int64_t per, double mean,
                                  double sd, int affine, hipStream_t s);
hipError_t launch_fill_exponential_f64(EngineState* st, int64_t n, double* out, int64_t per,
                                       double lambda, hipStream_t s);
hipError_t launch_fill_cauchy_f64(EngineState* st, int64_t n, double* out, int64_t per, double a,
                                  double b, hipStream_t s);
// NormalVGen; a NULL mean or sd array stands for mean0 or sd0 at every element.
hipError_t launch_fill_normal_vop_f32(EngineState* st, int64_t n, float* out, int64_t per,
                                      const float* mean, const float* sd, hipStream_t s, float mean0 = 0.f,
                                      float sd0 = 1.f);
hipError_t launch_fill_exponential_f32(EngineState* st, int64_t n, float* out, int64_t per,
                                       float lambda, hipStream_t s);
hipError_t launch_fill_cauchy_f32(EngineState* st, int64_t n, float* out, int64_t per, float a,
                                  float b, hipStream_t s);
hipError_t launch_fill_studentt_f32(EngineState* st, int64_t n, float* out, int64_t per, float nu,
                                    hipStream_t s);
hipError_t launch_fill_lognormal_f32(EngineState* st, int64_t n, float* out, int64_t per, float mean,
                                     float sd, hipStream_t s);
hipError_t launch_fill_weibull_f32(EngineState* st, int64_t n, float* out, int64_t per, float a, float b,
                                   hipStream_t s);
hipError_t launch_fill_extreme_value_f32(EngineState* st, int64_t n, float* out, int64_t per, float a,
                                         float b, hipStream_t s);
hipError_t launch_fill_gamma_fixed_f32(EngineState* st, int64_t n, float* out, int64_t per,
                                       float alpha, float beta, hipStream_t s);
// BernoulliGen<float> (p = uint32(p * 2^31) as int32) or BernoulliVGen (vp != NULL).
hipError_t launch_fill_bernoulli_f32(EngineState* st, int64_t n, float* out, int64_t per, int32_t p,
                                     const float* vp, hipStream_t s);
// Balanced2Gen(a, b), or BalancedVGen with per-element a, b (va != NULL).
hipError_t launch_fill_balanced_f32(EngineState* st, int64_t n, float* out, int64_t per, float a, float b,
                                    const float* va, const float* vb, hipStream_t s);
// Vectorize-over-params float fills with one or two per-element parameters.
enum VopKind : int {
    VOP_UNIFORM = 0,        // a, b
    VOP_EXPONENTIAL = 1,    // lambda
    VOP_CAUCHY = 2,         // a, b
    VOP_WEIBULL = 3,        // a, b
    VOP_EXTREME_VALUE = 4,  // a, b
    VOP_STUDENT_T = 5,      // n
    VOP_LOGNORMAL = 6,      // mean, sd
};
hipError_t launch_fill_vop_f32(EngineState* st, int64_t n, int kind, float* out, int64_t per, const float* va,
                               const float* vb, hipStream_t s);
// Half-draw int32 fills (Packet4i through half()).
hipError_t launch_fill_bernoulli_i32(EngineState* st, int64_t n, int32_t* out, int64_t per, int32_t p,
                                     hipStream_t s);
hipError_t launch_fill_geometric_i32(EngineState* st, int64_t n, int32_t* out, int64_t per, float rlog_q,
                                     hipStream_t s);
hipError_t launch_fill_randbits_i32(EngineState* st, int64_t n, int32_t* out, int64_t per, hipStream_t s);
// DiscreteGen<int32_t, double>: a double cdf (ncdf = K < 16) or the double alias table.
hipError_t launch_fill_discrete_d_i32(EngineState* st, int64_t n, int32_t* out, int64_t per, const double* cdf,
                                      int32_t ncdf, const double* prob, const int32_t* alias, uint32_t bitmask,
                                      hipStream_t s);
// DiscreteGen<int32_t, float>: cdf (k > 0, alias NULL) or float prob + alias tables.
hipError_t launch_fill_discrete_f_i32(EngineState* st, int64_t n, int32_t* out, int64_t per, const float* tab,
                                      const int32_t* alias, int32_t k, uint32_t bitmask, hipStream_t s);
// UniformIntGen<int32_t>(lo, lo + pdiff), pdiff < 2^31.
hipError_t launch_fill_uniform_int_i32(EngineState* st, int64_t n, int32_t* out, int64_t per, int32_t lo,
                                       uint32_t pdiff, hipStream_t s);

// fill_reject.hip (CompressMask rejection samplers); gs: optional per-engine
// queues of a reused generator object.
hipError_t launch_fill_gamma_reject_f32(EngineState* st, int64_t n, float* out, int64_t per,
                                        float alpha, float beta, Guard gd, hipStream_t s,
                                        GenState* gs = nullptr);
hipError_t launch_fill_poisson_reject_i32(EngineState* st, int64_t n, int32_t* out, int64_t per,
                                          const PoissonParams& pp, Guard gd, hipStream_t s,
                                          GenState* gs = nullptr);
hipError_t launch_fill_binomial_reject_i32(EngineState* st, int64_t n, int32_t* out, int64_t per,
                                           const BinomialParams& bp, Guard gd, hipStream_t s,
                                           GenState* gs = nullptr);
// Reused StdNormalGen / NormalGen (mode 0 / 1), NormalVGen (mode 2) or
// LognormalGen (mode 3):
// starts with the cached cos packet of gs and leaves a trailing one in it.
hipError_t launch_fill_normal_gen_f32(EngineState* st, int64_t n, float* out, int64_t per, int mode,
                                      float mean, float sd, const float* vmean, const float* vsd,
                                      GenState* gs, hipStream_t s);

// fill_lockstep.hip (packet-lockstep samplers on half draws)
hipError_t launch_fill_poisson_knuth_i32(EngineState* st, int64_t n, int32_t* out, int64_t per,
                                         double mean, int32_t trials, int flip, Guard gd,
                                         hipStream_t s);
hipError_t launch_fill_binomial_small_i32(EngineState* st, int64_t n, int32_t* out, int64_t per,
                                          int32_t trials, double p, hipStream_t s);
hipError_t launch_fill_binomial_vop_i32(EngineState* st, int64_t n, int32_t* out, int64_t per,
                                        Param<int32_t> trials, Param<float> p, Guard gd,
                                        hipStream_t s);
hipError_t launch_fill_binomial_vop_f32(EngineState* st, int64_t n, int32_t* out, int64_t per,
                                        Param<int32_t> trials, Param<float> p, Guard gd,
                                        hipStream_t s);

hipError_t launch_fill_gamma_vop_f32(EngineState* st, int64_t n, float* out, int64_t per,
                                     const float* alpha, const float* beta, Guard gd,
                                     hipStream_t s);
hipError_t launch_fill_poisson_vop_i32(EngineState* st, int64_t n, int32_t* out, int64_t per,
                                       const float* mean, Guard gd, hipStream_t s);

// program.hip (segment programs)
// Device table of plgamma_approx(k), k = 1..kLgTab, for the current device:
// built on first use (one kernel on `s`, then a synchronisation of `s`) and
// kept until the process ends. NULL if it cannot be allocated.
const float* lgamma_table(hipStream_t s);
// The current device's table if it has been built, else NULL (builds nothing).
const float* lgamma_table_built();
hipError_t launch_program(EngineState* st, int32_t n_groups, const int32_t* d_engine_of,
                          const int32_t* d_begin, const Segment* d_segs, Guard gd,
                          hipStream_t s);
// The same on P8 adaptors (group b = adaptor engine_of[b], records 2a, 2a+1);
// mode_override >= 0 forces the scalar packing, else each kind's own.
hipError_t launch_program_p8(EngineState* st, int32_t n_groups, const int32_t* d_engine_of, const int32_t* d_begin,
                             const Segment* d_segs, int mode_override, Guard gd, hipStream_t s);
Segment seg_poisson(int32_t* out, int64_t count, double mean);
Segment seg_binomial(int32_t* out, int64_t count, int32_t trials, double p);

}  // namespace erd
