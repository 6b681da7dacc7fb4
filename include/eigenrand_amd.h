/*
 * eigenrand_amd: C ABI of the MI355X (gfx950) implementation of EigenRand's
 * sampling hot path.
 *
 * The reference exposes this path as C++ templates that Eigen's assignment
 * loop drives packet by packet; this header is the flat, C-callable boundary a
 * binding (ctypes, cgo, JNI, a C++ shim) would target instead. Every symbol is
 * `extern "C"`, takes plain pointers and sizes, and returns an erd_status.
 *
 * Conventions
 *  - An erd_engines handle is a set of N Vmt19937_64 engines resident in
 *    device memory (reference: EigenRand/PacketRandomEngine.h:585-599, AVX2
 *    layout: each engine = 4 MT19937-64 lanes, lane l seeded seed+l).
 *  - A fill with `per_engine` outputs writes N contiguous blocks: engine g
 *    fills out[g*per_engine, (g+1)*per_engine) in the reference's packet order,
 *    exactly as `x = Rand::xxx<ArrayXf>(per_engine, 1, engine_g, ...)` would.
 *    Each call behaves like a free function of the reference (a fresh
 *    generator per engine block: queued rejection leftovers and the normal
 *    generator's cached half-packet are dropped at the end of the block), and
 *    advances the engines exactly as the reference advances them.
 *  - Output/parameter pointers are DEVICE pointers; `stream` is a hipStream_t
 *    (NULL = default stream). Calls are asynchronous on that stream.
 *  - per_engine must be a multiple of the packet width: 8 for float fills
 *    (Packet8f), 4 for int32/double/u64 fills. The reference's scalar tail
 *    path for other sizes is not part of this boundary (ERD_ERR_INVALID).
 */
#ifndef EIGENRAND_AMD_H
#define EIGENRAND_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum erd_status {
    ERD_OK = 0,
    ERD_ERR_INVALID = 1,     /* bad argument (size, pointer, parameter range)          */
    ERD_ERR_NONFINITE = 2,   /* NaN/Inf weights (reference: Discrete.h:88 throws)       */
    ERD_ERR_NOT_PSD = 3,     /* covariance not positive semi-definite (MvNormal.h:31-35) */
    ERD_ERR_ITER_CAP = 4,    /* a rejection loop hit the device iteration cap           */
    ERD_ERR_OOM = 5,         /* device allocation failed                                */
    ERD_ERR_HIP = 6,         /* a HIP runtime call failed                               */
    ERD_ERR_UNSUPPORTED = 7  /* parameter combination outside this boundary             */
} erd_status;

typedef struct erd_engines erd_engines;

/* ---------------------------------------------------------------- misc --- */
const char* erd_status_string(int status);
size_t erd_engine_state_bytes(void); /* bytes of one engine record (10016) */

/* ------------------------------------------------------------- engines --- */
/* Vmt19937_64(seed) for each engine: MersenneTwister(uint64_t x0)
 * (PacketRandomEngine.h:230-239). `seeds` is a HOST array of n_engines. */
int erd_engines_create(erd_engines** out, const uint64_t* seeds, int64_t n_engines, void* stream);
/* Engine g seeded seed0 + g*stride (stride 4 gives distinct lanes seed0.. ). */
int erd_engines_create_range(erd_engines** out, uint64_t seed0, uint64_t stride, int64_t n_engines,
                             void* stream);
/* Non-owning handle on engines [first, first+n) of `parent`. */
int erd_engines_view(erd_engines** out, erd_engines* parent, int64_t first, int64_t n);
int erd_engines_destroy(erd_engines* e);
int64_t erd_engines_count(const erd_engines* e);
/* Device address of the engine records (n * erd_engine_state_bytes()). */
void* erd_engines_device_ptr(erd_engines* e);
/* Copy engine records to/from HOST memory (the copy semantics of
 * AlignedArray / `auto copy = engine;`, PacketRandomEngine.h:89-118). */
int erd_engines_snapshot(const erd_engines* e, void* host_dst, void* stream);
int erd_engines_restore(erd_engines* e, const void* host_src, void* stream);
/* dst engines := src engines (same count), device to device. */
int erd_engines_copy(erd_engines* dst, const erd_engines* src, void* stream);
/* MersenneTwister::discard(calls) (PacketRandomEngine.h:317-323) on every engine. */
int erd_engines_discard(erd_engines* e, int64_t calls, void* stream);
/* Read and clear the device status bits raised by kernels on this set (sync):
 * ERD_STATUS_ITER_CAP when a sampler gave up at an iteration cap (the
 * outputs of that launch are then unspecified), ERD_STATUS_BAD_SEGMENT for a
 * malformed request program. Fills return ERD_OK when enqueued; this call is
 * how a caller learns that one of them gave up. */
#define ERD_STATUS_ITER_CAP 1u
#define ERD_STATUS_BAD_SEGMENT 2u
int erd_engines_status(erd_engines* e, uint32_t* status_out, void* stream);
/* Lane layout of the fills on this set: the float packet width the
 * reference's build takes from its ISA (views created from the set later
 * copy it).
 *   8  AVX2, the default: Vmt19937_64 = Pmt19937_64<Packet8i>
 *      (PacketRandomEngine.h:585-586), a call is 4 u64 lanes, Packet8f.
 *   4  SSE2 / NEON: Vmt19937_64 = Pmt19937_64<Packet4i>
 *      (PacketRandomEngine.h:587-588), a 2-lane engine seeded x0, x0+1,
 *      which is u64 lanes 0-1 of this set's engines (lanes 2-3 are stepped
 *      and not read); Packet4f / Packet2d of one call
 *      (arch/SSE/RandUtils.h:70-115).
 *  16  AVX-512: the AVX2 engine, a Packet16f / Packet8d is two calls with the
 *      first in the low half (arch/AVX512/RandUtils.h:37-46, evaluated left
 *      to right as the reference's default Clang preset does).
 * Fills with a form at every layout: erd_fill_u64, erd_fill_uniform_f32/_f64,
 * erd_fill_std_normal_f32/_f64, erd_fill_normal_f32/_f64,
 * erd_fill_lognormal_f32, erd_fill_exponential_f32/_f64,
 * erd_fill_cauchy_f32/_f64, erd_fill_weibull_f32, erd_fill_extreme_value_f32
 * and erd_fill_studentt_f32; per_engine is then a multiple of the layout's
 * packet width. Every other call that consumes draws returns
 * ERD_ERR_UNSUPPORTED on a set whose layout is not 8. Engine records,
 * snapshots, copies and discard do not depend on the layout. */
int erd_engines_set_layout(erd_engines* e, int layout);
int erd_engines_layout(const erd_engines* e); /* 0 for an invalid handle */
/* Iteration caps of the launches on this set (views created from it later
 * copy them). The reference's rejection and lockstep loops run until they
 * succeed (RandUtils.h:342-346; Discrete.h:860-873, 1139-1185); here a
 * rejection fill gives up after reject_cap candidates without an output
 * packet, and a lockstep packet after lockstep_cap iterations, and the launch
 * raises ERD_STATUS_ITER_CAP. The caps are checked once per wave step (16
 * candidates of a rejection fill; 1 to 16 iterations of a lockstep packet,
 * by sampler; the windowed Knuth fills finish a packet inside its window of
 * at most 64 slots and check the cap only for a packet that runs past it),
 * so a packet may run up to one step past its cap. 0 selects the default
 * (2^22 and 2^20); negative values are invalid. */
int erd_engines_set_iteration_caps(erd_engines* e, int32_t reject_cap, int32_t lockstep_cap);

/* ---------------------------------------------------- raw engine draws --- */
/* Vmt19937_64::operator() (PacketRandomEngine.h:295-310): out[4k+l] = lane l
 * of draw k; per_engine = 4 * draws. */
int erd_fill_u64(erd_engines* e, uint64_t* out, int64_t per_engine, void* stream);

/* ------------------------------------------------------- P8 adaptor --- */
/* P8_mt19937_64 = ParallelRandomEngineAdaptor<uint64_t, Vmt19937_64, 8>
 * (PacketRandomEngine.h:429-517, 600-609): adaptor k = the Vmt pair
 * (seeds[k], seeds[k]+4), so its 8 u64 lanes are mt19937_64(seed .. seed+7)
 * on any ISA: its scalar stream is out[8r + j] = mt19937_64(seed + j)[r]. A
 * set of n adaptors is an erd_engines set of 2n Vmt engines (adaptor k =
 * engines 2k, 2k+1); snapshot/copy/discard work on it as usual. The record of
 * engine 2k keeps the position in the operator() buffer (scalars of the last
 * refill already returned) in its last word. */
int erd_engines_create_p8(erd_engines** out, const uint64_t* seeds, int64_t n_p8, void* stream);
/* operator() of P8_mt19937_64, continuing from the buffer position:
 * per_p8 scalars (a multiple of 4) per adaptor. */
int erd_p8_fill_u64(erd_engines* e, uint64_t* out, int64_t per_p8, void* stream);
/* operator() of P8_mt19937_64_32: 16 uint32 per refill, Vmt i's 4 lanes as
 * low/high word pairs at [8i, 8i+8); per_p8 is a multiple of 16. Whole
 * refills: it starts with a fresh refill, dropping any scalars an earlier u64
 * or distribution fill left in the operator() buffer, and leaves the buffer
 * position at 0. */
int erd_p8_fill_u32(erd_engines* e, uint32_t* out, int64_t per_p8, void* stream);
/* uniform_real() of the adaptor: bit_to_ur_float of the same 16 words per
 * refill; per_p8 is a multiple of 16; whole refills as above. */
int erd_p8_fill_uniform_f32(erd_engines* e, float* out, int64_t per_p8, void* stream);
/* A distribution fill with P8_mt19937_64 as its engine, as the reference's
 * own tests draw (test/test.cpp:117 and on): `Rand::xxx<ArrayXf>(per_p8, 1,
 * p8_k, a, b)` for each adaptor k, out = n blocks of per_p8. The packet code
 * takes the adaptor as a scalar full-bit engine and packs its scalars with
 * RawbitsMaker<..., scalar_fullbit> (arch/AVX/RandUtils.h:59-144,
 * arch/SSE/RandUtils.h:22-67). Those packings call _mm*_set*_epi64x(rng(),
 * rng(), ...), whose argument evaluation order C++ leaves unspecified; this
 * follows Clang's left-to-right order (the reference's default toolchain
 * preset, CMakePresets.json:16-17):
 *  - float samplers (Packet8f uniform_real, rawbits_34): three scalars a, b, c
 *    per draw give the 32-bit words [a.lo, a.hi, b.lo, X1, b.hi, c.lo, c.hi,
 *    X2], X1 = the top bytes of a.lo, a.hi, b.lo, b.lo, X2 those of b.hi,
 *    c.lo, c.hi, c.hi. dist: uniform, normal (std or mean/stdev), exponential,
 *    Cauchy, Student-t, gamma, lognormal, Weibull, extreme value (float out);
 *  - int32 samplers (Packet4i rawbits, _mm_set_epi64x(rng(), rng())): two
 *    scalars per half draw, the second in u64 lane 0. dist: Poisson,
 *    binomial (int32 out).
 * Other kinds return ERD_ERR_UNSUPPORTED. a, b as in erd_request; per_p8 a
 * multiple of the packet width (8 floats, 4 int32). */
int erd_p8_fill_dist(erd_engines* e, int32_t dist, void* out, int64_t per_p8, double a, double b, void* stream);

/* ---------------------------------------------------------- univariate --- */
/* uniformReal<float>: StdUniformRealGen (Basic.h:584-590); with (a,b) it is
 * UniformRealGen (Basic.h:626-634): u*(b-a)+a. Identical values for (0,1). */
int erd_fill_uniform_f32(erd_engines* e, float* out, int64_t per_engine, float a, float b,
                         void* stream);
/* uniformReal<double> (arch/AVX/RandUtils.h:212-225): 4 doubles per draw. */
int erd_fill_uniform_f64(erd_engines* e, double* out, int64_t per_engine, double a, double b,
                         void* stream);
/* f64 forms at Packet4d (4 doubles per engine draw, per_engine a multiple of
 * 4). Box-Muller, the log and the sincos follow the reference's double packet
 * code: plog<Packet4d> (arch/AVX/MorePacketMath.h:781-858) and the double
 * psincos (MorePacketMath.h:396-491). normal<double>: StdNormalGen
 * (NormalExp.h:60-82), pairs of draws give [r sin x4, r cos x4]; with
 * (mean, stdev) it is NormalGen (NormalExp.h:127-135). A per_engine that is 4
 * mod 8 ends with a sine packet whose cosine packet is dropped, as a fresh
 * generator per call does. */
int erd_fill_std_normal_f64(erd_engines* e, double* out, int64_t per_engine, void* stream);
int erd_fill_normal_f64(erd_engines* e, double* out, int64_t per_engine, double mean, double stdev,
                        void* stream);
/* exponential<double>(…, lambda): ExponentialGen (NormalExp.h:391-398). */
int erd_fill_exponential_f64(erd_engines* e, double* out, int64_t per_engine, double lambda,
                             void* stream);
/* cauchy<double>(…, a, b): CauchyGen (NormalExp.h:843-855). */
int erd_fill_cauchy_f64(erd_engines* e, double* out, int64_t per_engine, double a, double b,
                        void* stream);
/* normal<float>(rows, cols, urng): StdNormalGen (NormalExp.h:60-82). */
int erd_fill_std_normal_f32(erd_engines* e, float* out, int64_t per_engine, void* stream);
/* normal<float>(rows, cols, urng, mean, stdev): NormalGen (NormalExp.h:127-135). */
int erd_fill_normal_f32(erd_engines* e, float* out, int64_t per_engine, float mean, float stdev,
                        void* stream);
/* normal(urng, mean_array, stdev_array): NormalVGen (NormalExp.h:138-163,
 * free function 1294-1335). mean/stdev are device arrays of n*per_engine. */
int erd_fill_normal_vop_f32(erd_engines* e, float* out, int64_t per_engine, const float* mean,
                            const float* stdev, void* stream);
/* normal<float> in the FAST precision mode (StdNormalGen / NormalGen,
 * NormalExp.h:60-82, 127-135, with other arithmetic): the same engine draws,
 * the same Box-Muller pairing and output order, and the same final engine
 * state as erd_fill_std_normal_f32 / erd_fill_normal_f32, but r and the angle
 * use the hardware log2, square root, sine and cosine instead of the
 * reference's cephes plog and psincos. Each output z stays within
 *   |z - z_exact| <= 4 ulp(z_exact) + 2^-20 * r,   r = sqrt(-2 ln u1) <= 5.65,
 * of the bit-exact mode's value (asserted by the tests). */
int erd_fill_std_normal_f32_fast(erd_engines* e, float* out, int64_t per_engine, void* stream);
int erd_fill_normal_f32_fast(erd_engines* e, float* out, int64_t per_engine, float mean, float stdev,
                             void* stream);
/* normal(urng, mean_array, stdev) and normal(urng, mean, stdev_array): the
 * reference's mixed overloads (NormalExp.h:1304-1335), NormalVGen with the
 * scalar as a constant array of the same shape, so each equals
 * erd_fill_normal_vop_f32 with that array. */
int erd_fill_normal_vop_vs_f32(erd_engines* e, float* out, int64_t per_engine, const float* mean, float stdev,
                               void* stream);
int erd_fill_normal_vop_sv_f32(erd_engines* e, float* out, int64_t per_engine, float mean, const float* stdev,
                               void* stream);
/* lognormal<float>(…, mean, stdev): LognormalGen (NormalExp.h:171-210), the
 * exp of NormalGen's packets (same cos-packet caching). */
int erd_fill_lognormal_f32(erd_engines* e, float* out, int64_t per_engine, float mean, float stdev,
                           void* stream);
/* weibull<float>(…, a, b): WeibullGen (NormalExp.h:614-656). */
int erd_fill_weibull_f32(erd_engines* e, float* out, int64_t per_engine, float a, float b, void* stream);
/* extremeValue<float>(…, a, b): ExtremeValueGen (NormalExp.h:690-732). */
int erd_fill_extreme_value_f32(erd_engines* e, float* out, int64_t per_engine, float a, float b,
                               void* stream);
/* exponential<float>(…, lambda): ExponentialGen (NormalExp.h:356-399). */
int erd_fill_exponential_f32(erd_engines* e, float* out, int64_t per_engine, float lambda,
                             void* stream);
/* cauchy<float>(…, a, b): CauchyGen (NormalExp.h:809-856). */
int erd_fill_cauchy_f32(erd_engines* e, float* out, int64_t per_engine, float a, float b,
                        void* stream);
/* studentT<float>(…, n): StudentTGen (NormalExp.h:241-302). */
int erd_fill_studentt_f32(erd_engines* e, float* out, int64_t per_engine, float n, void* stream);
/* gamma<float>(…, alpha, beta): GammaGen (NormalExp.h:434-606), all regimes. */
int erd_fill_gamma_f32(erd_engines* e, float* out, int64_t per_engine, float alpha, float beta,
                       void* stream);
/* chiSquared<float>(…, n) = GammaGen(n*0.5, 2) (NormalExp.h:767-802). */
int erd_fill_chisquared_f32(erd_engines* e, float* out, int64_t per_engine, float n, void* stream);

/* bernoulli<float>(…, p): BernoulliGen (Basic.h:670-708) at Packet8f, one draw
 * per 8 outputs of 1.0f/0.0f; 0 <= p <= 1. The packet path compares 31 random
 * bits with uint32(p * 2^31) as signed int32 lanes, so p = 1 gives zeros. */
int erd_fill_bernoulli_f32(erd_engines* e, float* out, int64_t per_engine, double p, void* stream);
/* bernoulli(urng, p_array): BernoulliVGen<float> (Basic.h:711-734). */
int erd_fill_bernoulli_vop_f32(erd_engines* e, float* out, int64_t per_engine, const float* p,
                               void* stream);
/* balanced<float>(…) = (a, b) = (-1, 1): BalancedGen (Basic.h:408-430);
 * balanced<float>(…, a, b): Balanced2Gen (Basic.h:437-470), zero_to_one *
 * (b - a) + a with zero_to_one = (w & 0x7FFFFFFF) / 0x7FFFFFFF in float. */
int erd_fill_balanced_f32(erd_engines* e, float* out, int64_t per_engine, float a, float b, void* stream);
/* balanced(urng, a_array, b_array): BalancedVGen (Basic.h:472-495). */
int erd_fill_balanced_vop_f32(erd_engines* e, float* out, int64_t per_engine, const float* a,
                              const float* b, void* stream);
/* Vectorize-over-params forms of the single-draw generators; parameters are
 * device arrays of n*per_engine elements laid out like the output.
 *   uniformReal(urng, a, b):      UniformRealVGen  (Basic.h:638-661)
 *   exponential(urng, lambda):    ExponentialVGen  (NormalExp.h:402-425)
 *   cauchy(urng, a, b):           CauchyVGen       (NormalExp.h:859-885)
 *   lognormal(urng, mean, stdev): LognormalVGen    (NormalExp.h:213-234), pexp of NormalVGen
 *   studentT(urng, n):            StudentTVGen     (NormalExp.h:305-346)
 *   weibull(urng, a, b):          WeibullVGen      (NormalExp.h:659-682)
 *   extremeValue(urng, a, b):     ExtremeValueVGen (NormalExp.h:736-761) */
int erd_fill_uniform_vop_f32(erd_engines* e, float* out, int64_t per_engine, const float* a, const float* b,
                             void* stream);
int erd_fill_exponential_vop_f32(erd_engines* e, float* out, int64_t per_engine, const float* lambda,
                                 void* stream);
int erd_fill_cauchy_vop_f32(erd_engines* e, float* out, int64_t per_engine, const float* a, const float* b,
                            void* stream);
int erd_fill_lognormal_vop_f32(erd_engines* e, float* out, int64_t per_engine, const float* mean,
                               const float* stdev, void* stream);
int erd_fill_studentt_vop_f32(erd_engines* e, float* out, int64_t per_engine, const float* n, void* stream);
int erd_fill_weibull_vop_f32(erd_engines* e, float* out, int64_t per_engine, const float* a, const float* b,
                             void* stream);
int erd_fill_extreme_value_vop_f32(erd_engines* e, float* out, int64_t per_engine, const float* a,
                                   const float* b, void* stream);

/* beta<float>(…, a, b): BetaGen (NormalExp.h:896-985). a < 1 and b < 1:
 * Johnk's rejection through CompressMask (2 draws per candidate); otherwise
 * per packet GammaGen(a) then GammaGen(b) (each with its own rejection queue),
 * p1 / (p1 + p2). a, b > 0. */
int erd_fill_beta_f32(erd_engines* e, float* out, int64_t per_engine, float a, float b, void* stream);
/* fisherF<float>(…, m, n): FisherFGen (NormalExp.h:987-1023) = BetaGen(m/2,
 * n/2) mapped by (b/a) x / (1 - x). m, n > 0. */
int erd_fill_fisher_f_f32(erd_engines* e, float* out, int64_t per_engine, float m, float n, void* stream);

/* ------------------------------------------------------ integer (int32) --- */
/* bernoulli<int32>(…, p): BernoulliGen (Basic.h:670-708) at Packet4i, one
 * half() draw per 4 outputs of 1/0; 0 <= p <= 1. */
int erd_fill_bernoulli_i32(erd_engines* e, int32_t* out, int64_t per_engine, double p, void* stream);
/* geometric<int32>(…, p): GeometricGen (Discrete.h:1339-1381), trunc(log(1 -
 * U) * (1 / log(1 - p))) from a half() draw per 4 outputs, converted as
 * cvttps2dq does (out of range: INT32_MIN); 0 <= p <= 1. */
int erd_fill_geometric_i32(erd_engines* e, int32_t* out, int64_t per_engine, double p, void* stream);
/* negativeBinomial<int32>(…, trials, p): NegativeBinomialGen
 * (GammaPoisson.h:29-100) at Packet4i: per packet a GammaGen<float>(trials,
 * (1-p)/p) packet on half draws, then Poisson counts with those means (Knuth
 * if no mean exceeds 12, else tangent rejection sharing the gamma's
 * CompressMask queue, as the reference does). trials > 0, 0 < p <= 1. */
int erd_fill_negative_binomial_i32(erd_engines* e, int32_t* out, int64_t per_engine, int32_t trials, double p,
                                   void* stream);
/* uniformInt<int32>(…, min, max): UniformIntGen (Discrete.h:215-335) at
 * Packet4i: candidates of `bitsize` bits from half() draws through
 * CompressMask<16>, the rest of a draw dropped once a packet is emitted.
 * Requires 0 <= max - min < 2^31 (ERD_ERR_INVALID if max < min,
 * ERD_ERR_UNSUPPORTED beyond, where the reference's int32 max - min overflows). */
int erd_fill_uniform_int_i32(erd_engines* e, int32_t* out, int64_t per_engine, int32_t min, int32_t max,
                             void* stream);
/* randBits<int32>(…): RandbitsGen (Basic.h:378-401) at Packet4i, the bits of
 * one half() draw per 4 outputs. */
int erd_fill_randbits_i32(erd_engines* e, int32_t* out, int64_t per_engine, void* stream);
/* poisson<int32>(…, mean): PoissonGen (Discrete.h:788-904). mean >= 0. */
int erd_fill_poisson_i32(erd_engines* e, int32_t* out, int64_t per_engine, double mean, void* stream);
/* binomial<int32>(…, trials, p): BinomialGen (Discrete.h:911-1053). */
int erd_fill_binomial_i32(erd_engines* e, int32_t* out, int64_t per_engine, int32_t trials, double p,
                          void* stream);
/* binomial(urng, trials_array, p_array): BinomialVGen<int32_t> (Discrete.h:
 * 1061-1193, free function 1805-1831). trials (int32) and p (float) are device
 * arrays of n*per_engine elements. */
int erd_fill_binomial_vop_i32(erd_engines* e, int32_t* out, int64_t per_engine, const int32_t* trials,
                              const float* p, void* stream);
/* binomial(urng, trials_array, p) and binomial(urng, trials, p_array): the
 * mixed overloads (Discrete.h:1852-1897), BinomialVGen<int32_t> with the
 * scalar as a constant array; each equals erd_fill_binomial_vop_i32 with that
 * array. */
int erd_fill_binomial_vop_vs_i32(erd_engines* e, int32_t* out, int64_t per_engine, const int32_t* trials, float p,
                                 void* stream);
int erd_fill_binomial_vop_sv_i32(erd_engines* e, int32_t* out, int64_t per_engine, int32_t trials, const float* p,
                                 void* stream);
/* binomial(urng, trials, p) with float-typed trials: BinomialVGen<float>
 * (Discrete.h:1196-1331; free functions 1824-1848 and the scalar-p form
 * 1851-1881 that MultinomialGen uses). Packet8f lanes on full draws, so the
 * stream differs from the int32 form's. The reference carries trials and
 * results as int32 bit patterns in float lanes; here they are int32 arrays.
 * per_engine is a multiple of 8. */
int erd_fill_binomial_vop_f32(erd_engines* e, int32_t* out, int64_t per_engine, const int32_t* trials,
                              const float* p, void* stream);
int erd_fill_binomial_vop_vs_f32(erd_engines* e, int32_t* out, int64_t per_engine, const int32_t* trials, float p,
                                 void* stream);

/* Vectorize-over-params gamma and Poisson. The reference has no GammaVGen or
 * PoissonVGen (its VoP generators stop at BinomialVGen); these follow the
 * BinomialVGen design (Discrete.h:1124-1192): per-element parameters in float,
 * lockstep candidate iterations, each lane keeps its first accepted candidate.
 * Gamma: Packet8f, 3 full draws per iteration, GS step for alpha <= 1, tangent
 * step otherwise. Poisson: Packet4f via half draws, Knuth lockstep for
 * mean < 12, tangent lockstep otherwise. Device arrays of n*per_engine. */
int erd_fill_gamma_vop_f32(erd_engines* e, float* out, int64_t per_engine, const float* alpha,
                           const float* beta, void* stream);
int erd_fill_poisson_vop_i32(erd_engines* e, int32_t* out, int64_t per_engine, const float* mean,
                             void* stream);

/* ------------------------------------------------------ batched requests --- */
/* Heterogeneous fills in one launch. Requests naming the same engine run on
 * it in array order; different engines run concurrently. Each request gives
 * exactly the result of the corresponding erd_fill_* call on a one-engine view
 * of `engine` issued in that order (the reference equivalent: successive
 * free-function fills on one engine object). */
typedef enum erd_dist {
    ERD_DIST_U64 = 0,           /* count = 4*draws                           */
    ERD_DIST_UNIFORM_F32 = 1,   /* a=min, b=max                              */
    ERD_DIST_UNIFORM_F64 = 2,   /* a=min, b=max                              */
    ERD_DIST_STD_NORMAL_F32 = 3,
    ERD_DIST_NORMAL_F32 = 4,    /* a=mean, b=stdev                           */
    ERD_DIST_NORMAL_VOP_F32 = 5,/* in0=mean[], in1=stdev[]                   */
    ERD_DIST_EXPONENTIAL_F32 = 6, /* a=lambda                                */
    ERD_DIST_CAUCHY_F32 = 7,    /* a, b                                      */
    ERD_DIST_STUDENT_T_F32 = 8, /* a=n                                       */
    ERD_DIST_GAMMA_F32 = 9,     /* a=alpha, b=beta                           */
    ERD_DIST_POISSON_I32 = 10,  /* a=mean                                    */
    ERD_DIST_BINOMIAL_I32 = 11, /* a=trials, b=p                             */
    ERD_DIST_BINOMIAL_VOP_I32 = 12, /* in0=trials[] (int32), in1=p[] (float) */
    ERD_DIST_GAMMA_VOP_F32 = 13,    /* in0=alpha[], in1=beta[]               */
    ERD_DIST_POISSON_VOP_I32 = 14,  /* in0=mean[] (float)                    */
    ERD_DIST_DISCARD = 15,          /* count = draws skipped, out unused      */
    ERD_DIST_LOGNORMAL_F32 = 16,    /* a=mean, b=stdev of the log             */
    ERD_DIST_WEIBULL_F32 = 17,      /* a=shape, b=scale                       */
    ERD_DIST_EXTREME_VALUE_F32 = 18 /* a=location, b=scale                    */
} erd_dist;

typedef struct erd_request {
    int32_t dist;       /* erd_dist                                       */
    int32_t engine;     /* index into the engine set                      */
    int64_t count;      /* outputs (a multiple of the packet width)       */
    void* out;          /* device output                                  */
    const void* in0;    /* device per-element parameters (VoP kinds)      */
    const void* in1;
    double a, b;        /* scalar parameters                              */
} erd_request;

int erd_run_requests(erd_engines* e, const erd_request* requests, int64_t n_requests, void* stream);

/* No counterpart in the reference: erd_run_requests for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). Same requests, kernels and results, but
 * the segment program goes into a fixed device array of the library (about
 * 290 KB, held on each GPU it runs on until the process ends) instead of a
 * stream-ordered allocation, so at most ERD_NOALLOC_MAX_REQUESTS requests per
 * call (more: ERD_ERR_INVALID). It allocates nothing; the engine set must have
 * been created on the current device (its creation builds the table the
 * tangent samplers read; without it: ERD_ERR_UNSUPPORTED). Calls that overlap
 * on one GPU (other streams or threads) share the array and race; calls on
 * one stream are ordered. */
#define ERD_NOALLOC_MAX_REQUESTS 4096
int erd_run_requests_noalloc(erd_engines* e, const erd_request* requests, int64_t n_requests, void* stream);

/* ---------------------------------------------------- generator objects --- */
/* A reused generator object on n engines: `auto gen = Rand::makeXxxGen(a, b);`
 * followed by several gen.generate(per_engine, 1, engine_g) calls. The
 * object keeps, per engine, what the reference's generator keeps between
 * calls (Basic.h:268-342): StdNormalGen/NormalGen's cached cos packet
 * (NormalExp.h:64-68, 80) and the CompressMask leftovers of GammaGen,
 * PoissonGen and BinomialGen in their rejection regimes (NormalExp.h:440,
 * Discrete.h:793). Consecutive fills therefore equal one larger fill, while
 * the plain erd_fill_* calls behave like the free functions (a fresh
 * generator per call). dist: an erd_dist scalar kind with a, b as in
 * erd_request (uniform, normal, lognormal, exponential, Cauchy, Student-t,
 * Weibull, extreme value, gamma, Poisson, binomial). Fills must use an
 * engine set of n_engines engines. */
typedef struct erd_generator erd_generator;
int erd_generator_create(erd_generator** out, int32_t dist, double a, double b, int64_t n_engines,
                         void* stream);
int erd_generator_destroy(erd_generator* g);
int erd_generator_reset(erd_generator* g, void* stream);  /* empties the caches */
int erd_generator_fill(erd_generator* g, erd_engines* e, void* out, int64_t per_engine, void* stream);

/* ---------------------------------------------------- multivariate batch --- */
/* Batched FP64 Cholesky (the factorisation MvNormalGen/WishartGen take from
 * detail::get_lt, MvNormal.h:20-38, done in FP64 on the GPU instead of the
 * reference's float LLT). `a`: DEVICE, n_inst column-major d_i x d_i matrices
 * back to back, factored in place (L, upper part zeroed). dims/info: HOST.
 * Returns ERD_ERR_NOT_PSD if some info[i] != 0 (the reference then falls back
 * to an eigen-decomposition; this boundary reports it instead); the matrix of
 * an instance with info[i] != 0 is left partly factored. */
int erd_cholesky_f64_batch(int64_t n_inst, const int32_t* dims, double* a, int32_t* info, void* stream);

/* MvNormalGen<float>(mean_i, cov_i).generate(engine_i, samples_i) for
 * i < n_inst (MvNormal.h:56-116); instance i uses engine i of `e`.
 * dims/samples: HOST arrays; cov: DEVICE doubles (column-major d_i x d_i, back
 * to back); mean: DEVICE floats (d_i each, back to back) or NULL for zero;
 * out: DEVICE floats, X_i = L_i Z_i + mean_i (d_i x samples_i, column-major,
 * back to back). Z_i is the engine's StdNormalGen fill of d_i*samples_i
 * (bit-exact to the reference); optional z_out receives it. L_i comes from
 * the FP64 Cholesky; L*Z runs as split-BF16 (hi+lo) MFMA with FP32
 * accumulation. d_i*samples_i must be a multiple of 8. */
int erd_mvnormal_batch(erd_engines* e, int64_t n_inst, const int32_t* dims, const int64_t* samples,
                       const double* cov, const float* mean, float* out, float* z_out, void* stream);

/* Generator construction and generation as separate steps, as the reference
 * does (MvNormalGen's and WishartGen's constructors factorise once,
 * MvNormal.h:20-38, 86-98, 196-215; generate() reuses the factor).
 * erd_mv_factors_create factorises n_inst covariances (or Wishart scales;
 * DEVICE doubles as above, dims a HOST array) in FP64 on the device and keeps
 * the factors; info (HOST, optional, n_inst) receives each instance's
 * Cholesky status (0 = ok, k > 0: pivot k not positive), and a non-positive
 * pivot returns ERD_ERR_NOT_PSD. Generation uses instance i on engine i. */
typedef struct erd_mv_factors erd_mv_factors;
int erd_mv_factors_create(erd_mv_factors** out, int64_t n_inst, const int32_t* dims, const double* a,
                          int32_t* info, void* stream);
int erd_mv_factors_destroy(erd_mv_factors* f);
int64_t erd_mv_factors_count(const erd_mv_factors* f);
/* MvNormalGen::generate for every instance of f (arguments as in
 * erd_mvnormal_batch). */
int erd_mvnormal_generate(erd_engines* e, const erd_mv_factors* f, const int64_t* samples, const float* mean,
                          float* out, float* z_out, void* stream);
/* WishartGen::generate for every instance of f (arguments as in
 * erd_wishart_batch). */
int erd_wishart_generate(erd_engines* e, const erd_mv_factors* f, const int64_t* df, const int64_t* samples,
                         float* out, void* stream);

/* WishartGen<float>(df_i, scale_i).generate(engine_i, samples_i)
 * (MvNormal.h:184-302): out_i is d_i x (d_i*samples_i) column-major (sample j
 * in columns [j*d_i, (j+1)*d_i)). samples_i must be a multiple of 8 and
 * df_i > d_i - 1. scale: DEVICE doubles as for cov above. */
int erd_wishart_batch(erd_engines* e, int64_t n_inst, const int32_t* dims, const int64_t* df,
                      const int64_t* samples, const double* scale, float* out, void* stream);
/* InvWishartGen<float>(df_i, scale_i) (MvNormal.h:340-466): factors of the
 * inverse scale (get_lt(scale.inverse()), MvNormal.h:375-380, computed in FP64
 * on the device from the Cholesky factor of the scale; info and NOT_PSD refer
 * to that factorisation), then per sample the Wishart matrix T = L A, U =
 * T^-1 (lower triangular) and out = U^T U. Arguments and layout as for
 * erd_wishart_batch; d_i <= 512 (ERD_ERR_UNSUPPORTED beyond). */
int erd_mv_factors_create_inverse(erd_mv_factors** out, int64_t n_inst, const int32_t* dims, const double* scale,
                                  int32_t* info, void* stream);
int erd_invwishart_generate(erd_engines* e, const erd_mv_factors* f, const int64_t* df, const int64_t* samples,
                            float* out, void* stream);
int erd_invwishart_batch(erd_engines* e, int64_t n_inst, const int32_t* dims, const int64_t* df,
                         const int64_t* samples, const double* scale, float* out, void* stream);

/* DirichletGen<float>(alpha_i).generate(engine_i, samples_i)
 * (Multinomial.h:159-217): out_i is K_i x samples_i column-major. alpha: HOST
 * floats (K_i each, back to back); samples_i must be a multiple of 8. The
 * call may split its work over `stream` and an internal second stream of the
 * device (event fork and join): on return, everything it enqueued is ordered
 * before later work on `stream`. */
int erd_dirichlet_batch(erd_engines* e, int64_t n_inst, const int32_t* dims, const int64_t* samples,
                        const float* alpha, float* out, void* stream);

/* ------------------------------------------------- discrete / multinomial --- */
/* Host-side table of DiscreteGen<int32_t, int32_t> (Discrete.h:367-391): for
 * K < 16 the uint32 CDF (K entries) into cdf; otherwise AliasMethod
 * (Discrete.h:78-192) prob/alias of 2^ceil(log2 K) entries. HOST arrays.
 * ERD_ERR_NONFINITE if the weights do not sum to a finite value. */
int erd_alias_table(const double* weights, int64_t K, int32_t* prob, int32_t* alias, uint32_t* cdf);

/* discrete<int32>(…, weights): DiscreteGen<int32_t, int32_t> packet path
 * (Discrete.h:443-483, AVX2: a Packet8i per draw handed out as two Packet4i).
 * weights: HOST doubles; per_engine a multiple of 4. */
int erd_fill_discrete_i32(erd_engines* e, int32_t* out, int64_t per_engine, const double* weights, int64_t K,
                          void* stream);
/* discreteF<int32>(…, weights): DiscreteGen<int32_t, float> (Discrete.h:
 * 486-603) at Packet4i: K < 16 counts a half draw's uniform against the float
 * CDF; otherwise the alias table in float, index bits from one half draw and
 * the uniform from the next. weights: HOST doubles (K). */
int erd_fill_discrete_f_i32(erd_engines* e, int32_t* out, int64_t per_engine, const double* weights, int64_t K,
                            void* stream);
/* discreteD<int32>(…, weights): DiscreteGen<int32_t, double> (Discrete.h:
 * 606-779) at Packet4i under AVX (Eigen 3.x). K < 16: one full draw per
 * packet gives four doubles counted against the double CDF; otherwise the
 * alias table in double, index bits from one half() draw and the uniform
 * doubles from the next full draw (so two packets take three draws when the
 * engine's half cache starts empty). weights: HOST doubles (K). */
int erd_fill_discrete_d_i32(erd_engines* e, int32_t* out, int64_t per_engine, const double* weights, int64_t K,
                            void* stream);

/* MultinomialGen<int32_t>(trials_i, weights_i).generate(engine_i, samples_i)
 * (Multinomial.h:25-134): out_i is K_i x samples_i column-major int32.
 * dims/trials/samples/weights: HOST; out: DEVICE. trials < max(30(K-1),100)
 * uses per-sample alias/CDF draws, otherwise K-1 BinomialVGen<float> rows
 * (samples_i must then be a multiple of 8). */
int erd_multinomial_batch(erd_engines* e, int64_t n_inst, const int64_t* dims, const int32_t* trials,
                          const int64_t* samples, const double* weights, int32_t* out, void* stream);

/* ---------------------------------------------------------------- serving --- */
/* A request loop with continuous batching. Each client (any 64-bit id) owns
 * E = engines_per_client Vmt19937_64 engines resident on the device, engine g
 * seeded splitmix64(client ^ seed_salt) + 4g. A request spreads over
 * e = min(E, ceil(count / chunk)) consecutive engines of its client, starting
 * at a per-client cursor that advances by e (mod E) in submission order; chunk
 * gives each engine about 2^16 normal-equivalents of work (heavier samplers
 * spread wider), and each engine fills an equal, packet-aligned block
 * (MvNormal/multinomial: blocks of samples). Requests queue as they arrive; a
 * launcher thread packs the oldest admissible ones (none of their engines held
 * by a batch in flight, output within the HBM budget) into one launch on a
 * free stream while earlier batches run; a completion thread stamps them done.
 * A result depends only on the client's request sequence, never on what else
 * is in the batch. Outputs are allocated by the server and counted against
 * hbm_budget_bytes until erd_server_release. No counterpart in the reference
 * (which has no runtime). */
#define ERD_SERVE_MVNORMAL 100    /* dim=d, count=samples, host_param=cov (d*d col-major), host_mean */
#define ERD_SERVE_MULTINOMIAL 101 /* dim=K, count=samples, a=trials, host_param=weights (K)        */

typedef struct erd_server erd_server;

typedef struct erd_server_config {
    int64_t max_clients;
    int32_t engines_per_client;
    int32_t max_inflight;        /* batches running at once (one stream each) */
    int64_t hbm_budget_bytes;    /* admission cap on outputs not yet released */
    int64_t max_batch_requests;
    uint64_t seed_salt;
    int64_t max_batch_elements;  /* cap on one batch's outputs (0: none); a batch takes >= 1 request */
    int64_t engine_work;         /* normal-equivalent samples per engine when a request is spread
                                    over its client's engines (0: 65536) */
    /* Iteration caps of the batches (0: the defaults, as for
     * erd_engines_set_iteration_caps). A batch whose sampler gives up fails all
     * its requests with ERD_ERR_ITER_CAP and returns no output for them
     * (erd_server_wait gives out = NULL). */
    int32_t reject_cap;
    int32_t lockstep_cap;
} erd_server_config;

typedef struct erd_serve_request {
    uint64_t client;
    int32_t kind;                /* erd_dist (scalar-parameter kinds) or ERD_SERVE_* */
    int32_t dim;                 /* MvNormal: d (0 with a model: the model's)  */
    int64_t count;               /* outputs, or samples for ERD_SERVE_*       */
    double a, b;                 /* scalar parameters as in erd_request       */
    const double* host_param;    /* copied at submit                          */
    const float* host_mean;      /* MvNormal: overrides the model's mean      */
    uint64_t model;              /* MvNormal: a registered model, or 0 to pass
                                    the covariance in host_param               */
} erd_serve_request;

typedef struct erd_server_stats {
    int64_t completed, elements, batches;
    double p50_ms, p99_ms, max_ms, span_s;
    int64_t budget_in_use;
    double prep_s;               /* launcher time spent building and launching batches */
} erd_server_stats_t;

int erd_server_create(erd_server** out, const erd_server_config* cfg);
int erd_server_destroy(erd_server* srv);
/* Registers a client's MvNormal model (HOST covariance d x d column-major,
 * optional HOST mean); requests name it by *model. Its FP64 factorisation runs
 * once, on its first request (MvNormalGen's constructor, MvNormal.h:86-98). */
int erd_server_register_mvnormal(erd_server* srv, uint64_t client, int32_t d, const double* cov,
                                 const float* mean, uint64_t* model);
int erd_server_submit(erd_server* srv, const erd_serve_request* req, uint64_t* ticket);
/* Waits (timeout_s < 0: forever); returns the request's status and its
 * device output (out_elems elements: e blocks of the padded per-engine size;
 * for MvNormal/multinomial a dim x (padded samples) column-major matrix). */
int erd_server_wait(erd_server* srv, uint64_t ticket, double timeout_s, void** out, int64_t* out_elems);
int erd_server_release(erd_server* srv, uint64_t ticket);
int erd_server_stats(erd_server* srv, erd_server_stats_t* stats);
/* Clears the latency record and counters (e.g. after a warm-up pass). */
int erd_server_reset_stats(erd_server* srv);
/* Non-owning view of a client's engines (after its first request ran). */
int erd_server_client_engines(erd_server* srv, uint64_t client, erd_engines** view);

/* ---------------------------------------------------------- measurement --- */
/* No reference counterpart: writes 1.0f to n_blocks contiguous blocks of
 * per_block floats (a multiple of 4) with one 64-thread workgroup per block
 * and 16-byte stores, the store pattern of the fills without their
 * arithmetic. It is the practical HBM-write ceiling the fills are compared
 * with. */
int erd_store_ceiling_f32(float* out, int64_t n_blocks, int64_t per_block, void* stream);

/* -------------------------------------------------------- device memory --- */
/* Small helpers so callers without another GPU runtime can use the library. */
int erd_device_alloc(void** ptr, size_t bytes);
int erd_device_free(void* ptr);
int erd_memcpy_d2h(void* host_dst, const void* dev_src, size_t bytes, void* stream);
int erd_memcpy_h2d(void* dev_dst, const void* host_src, size_t bytes, void* stream);
int erd_stream_synchronize(void* stream);

#ifdef __cplusplus
}
#endif
#endif /* EIGENRAND_AMD_H */
