// Segment-program kernel: one workgroup (one wave) per engine group runs that
// engine's list of fills in order, with the engine resident in LDS for the
// whole list. Each segment reproduces the corresponding reference fill
// exactly (same ops as the single-fill kernels), so a program is equivalent
// to issuing its fills one after another on that engine.
#include <cmath>
#include <algorithm>
#include "erd_device.hpp"
#include "erd_dists.hpp"
#include "erd_kernels.hpp"
#include "erd_ops.hpp"
#include "erd_segment.hpp"

namespace erd {

constexpr int kBetaMaxCarry = 37;  // a + b - 1 draws of an integer-shape beta unit

union ProgramLds {
    uint64_t carry_fixed[18 * kLanes];  // GammaIntF32<19>: 18 straddling draws
    RejectLds<GammaGS> gs;
    RejectLds<GammaTangent> gt;
    RejectLds<BetaJohnk> gt_beta;
    RejectLds<GammaGSW<4>> gs4;
    RejectLds<GammaTangentW<4>> gt4;
    RejectLds<NegBinTangent> nbt;
    RejectLds<PoissonTangent> pt;
    RejectLds<BinomialTangent> bt;
    struct {
        uint64_t carry[2 * kLanes];
    } vop;
    // integer-shape beta: draws straddling a twist; small-n binomial: up to 23 half slots
    uint64_t carry_beta[kBetaMaxCarry * kLanes];
    struct {
        uint64_t carry[2 * kLanes];
        alignas(16) float u[kKnuthStage];
    } knuth;
};

template <int A>
__device__ void run_gamma_int(Eng& e, ProgramLds& s, float* out, int64_t n, float beta, int lane) {
    GammaIntF32<A> d{out, beta};
    op_fixed(e, s.carry_fixed, n, d, lane);
}

// gs: the generator's CompressMask queue carried across calls (BetaGen's two
// GammaGen members), or NULL for a fresh generator.
__device__ void run_gamma(Eng& e, ProgramLds& s, float* out, int64_t n, float alpha, float beta,
                          Guard gd, int lane, GenState* gs = nullptr) {
    if (alpha < 1.0f) {
        GammaGS pol{beta, (float)(kE / (alpha + kE)), 1 / alpha, alpha - 1};
        op_reject(e, s.gs, out, n, pol, gd, lane, gs);
        return;
    }
    if (alpha == 1.0f) {
        ExponentialF32 d{out, 1.0f, beta, 1};
        op_fixed(e, s.carry_fixed, n, d, lane);
        return;
    }
    const int count = (int)alpha;
    if ((float)count == alpha && count < 20) {
        switch (count) {
            case 2: run_gamma_int<2>(e, s, out, n, beta, lane); return;
            case 3: run_gamma_int<3>(e, s, out, n, beta, lane); return;
            case 4: run_gamma_int<4>(e, s, out, n, beta, lane); return;
            case 5: run_gamma_int<5>(e, s, out, n, beta, lane); return;
            case 6: run_gamma_int<6>(e, s, out, n, beta, lane); return;
            case 7: run_gamma_int<7>(e, s, out, n, beta, lane); return;
            case 8: run_gamma_int<8>(e, s, out, n, beta, lane); return;
            case 9: run_gamma_int<9>(e, s, out, n, beta, lane); return;
            case 10: run_gamma_int<10>(e, s, out, n, beta, lane); return;
            case 11: run_gamma_int<11>(e, s, out, n, beta, lane); return;
            case 12: run_gamma_int<12>(e, s, out, n, beta, lane); return;
            case 13: run_gamma_int<13>(e, s, out, n, beta, lane); return;
            case 14: run_gamma_int<14>(e, s, out, n, beta, lane); return;
            case 15: run_gamma_int<15>(e, s, out, n, beta, lane); return;
            case 16: run_gamma_int<16>(e, s, out, n, beta, lane); return;
            case 17: run_gamma_int<17>(e, s, out, n, beta, lane); return;
            case 18: run_gamma_int<18>(e, s, out, n, beta, lane); return;
            default: run_gamma_int<19>(e, s, out, n, beta, lane); return;
        }
    }
    GammaTangent pol{beta, sqrtf(2 * alpha - 1), alpha - 1};
    op_reject(e, s.gt, out, n, pol, gd, lane, gs);
}

// Per-segment state that must outlive the ops' LDS union: BetaGen's two
// gamma queues and the packet pair being combined; NegativeBinomialGen's
// queue, its packet of means and per-lane tangent constants.
struct BetaLds {
    GenState q[2];
    float x[2][8];
    float c[3][4];
};

// BetaGen::packetOp (NormalExp.h:950-984) and FisherFGen (1017-1022):
// a < 1 and b < 1 is the Johnk rejection through CompressMask; otherwise each
// packet is gd1.packetOp (GammaGen(a)) then gd2.packetOp (GammaGen(b)), each
// with its own queue, combined as p1 / (p1 + p2). Generators are fresh per
// fill: the queues start empty and their leftovers are dropped.
__device__ void run_beta(Eng& e, ProgramLds& s, BetaLds& bl, float* out, int64_t n, float a, float b, float ratio,
                         Guard gd, int lane) {
    if (a < 1.0f && b < 1.0f) {
        BetaJohnk pol{1 / a, 1 / b, ratio};
        op_reject(e, s.gt_beta, out, n, pol, gd, lane);
        return;
    }
    const int ia = (int)a, ib = (int)b;
    if ((float)ia == a && (float)ib == b && ia >= 1 && ib >= 1 && ia < 20 && ib < 20) {
        // both gammas take a fixed number of draws: packets are independent
        op_fixed_rt<kBetaMaxCarry>(e, s.carry_beta, n, ia + ib, BetaIntF32{out, ia, ib, ratio}, lane);
        return;
    }
    if (lane < 2) bl.q[lane].count = 0;
    __syncthreads();
    for (int64_t pk = 0; pk < n / 8; ++pk) {
        run_gamma(e, s, bl.x[0], 8, a, 1.0f, gd, lane, &bl.q[0]);
        __syncthreads();
        run_gamma(e, s, bl.x[1], 8, b, 1.0f, gd, lane, &bl.q[1]);
        __syncthreads();
        if (lane < 8) {
            const float p1 = bl.x[0][lane], p2 = bl.x[1][lane];
            const float x = p1 / (p1 + p2);
            out[8 * pk + lane] = ratio != 0.0f ? (ratio * x) / (1.0f - x) : x;
        }
        __syncthreads();
    }
}

// discard(n) on the resident engine: only the block position and twists.
__device__ void run_discard(Eng& e, int64_t calls, int lane) {
    int64_t end = e.idx + calls;
    if (end > kN) {
        const int64_t twists = (end - 1) / kN;
        for (int64_t t = 0; t < twists; ++t) twist_lds(e.mt, lane);
        e.twisted = true;
        end -= twists * kN;
    }
    e.idx = (int)end;
}

// GammaGen<float> at Packet4f with integer alpha < 20 or alpha == 1
// (NormalExp.h:561-578): alpha (or one) half draws per packet.
__device__ void gamma_fixed_w4(Eng& e, uint64_t* carry, float* dst4, float alpha, float beta, int lane) {
    HalfCursor cur = cursor_begin(e, carry, lane);
    const int j = lane & 3;
    float x;
    if (alpha == 1.0f) {
        cur.ensure(1, lane);
        const float u = ur_f32(cur.word32(0, j));
        cur.advance(1);
        x = beta * -(plog(1.0f - u) / 1.0f);
    } else {
        float yx = 1.0f;
        for (int c = 0; c < (int)alpha; ++c) {
            cur.ensure(1, lane);
            const float u = nzur_f32(cur.word32(0, j));
            cur.advance(1);
            yx = c == 0 ? u : yx * u;
        }
        x = -(beta * plog(yx));
    }
    if (lane < 4) dst4[lane] = x;
    cursor_end(e, cur);
}

// Knuth's product loop with a threshold per lane (NegativeBinomialGen,
// GammaPoisson.h:57-67); returns lane j's count.
__device__ int32_t knuth_packet_lanes(Eng& e, uint64_t* carry, const float* thr4, Guard gd, int lane) {
    HalfCursor cur = cursor_begin(e, carry, lane);
    const int j = lane & 3;
    const float thr = thr4[j];
    float val = 1.0f;
    int32_t res = 0;
    for (int it = 0;; ++it) {
        cur.ensure(1, lane);
        val = val * ur_f32(cur.word32(0, j));
        cur.advance(1);
        const bool c = thr < val;
        if (!(__ballot(c) & 0xFull)) break;
        res += c ? 1 : 0;
        if (it > gd.lockstep_cap) {
            if (lane == 0) atomicOr(gd.status, kStatusIterCap);
            break;
        }
    }
    cursor_end(e, cur);
    return res;
}

// NegativeBinomialGen<int32_t>::packetOp (GammaPoisson.h:52-97) at Packet4i:
// a packet of means from GammaGen<float>(trials, (1 - p) / p) at Packet4f;
// if no lane's mean exceeds 12, the per-lane Knuth loop; otherwise tangent
// rejection whose candidates share the gamma's CompressMask queue. A fresh
// generator per fill.
__device__ void run_negbin(Eng& e, ProgramLds& s, BetaLds& bl, int32_t* out, int64_t n, float alpha, float beta,
                           Guard gd, const float* lg, int lane) {
    GenState* q = &bl.q[0];
    if (lane == 0) q->count = 0;
    __syncthreads();
    const int count = (int)alpha;
    for (int64_t pk = 0; pk < n / 4; ++pk) {
        float* mean = bl.x[0];
        if (alpha < 1.0f) {
            GammaGSW<4> pol{beta, (float)(kE / (alpha + kE)), 1 / alpha, alpha - 1};
            op_reject(e, s.gs4, mean, 4, pol, gd, lane, q);
        } else if (alpha == 1.0f || ((float)count == alpha && count < 20)) {
            gamma_fixed_w4(e, s.vop.carry, mean, alpha, beta, lane);
        } else {
            GammaTangentW<4> pol{beta, sqrtf(2 * alpha - 1), alpha - 1};
            op_reject(e, s.gt4, mean, 4, pol, gd, lane, q);
        }
        __syncthreads();
        bool big = false;
#pragma unroll
        for (int j = 0; j < 4; ++j) big |= 12.0f < mean[j];
        if (!big) {
            if (lane < 4) bl.x[1][lane] = pexp(-mean[lane]);
            __syncthreads();
            const int32_t r = knuth_packet_lanes(e, s.vop.carry, bl.x[1], gd, lane);
            if (lane < 4) out[4 * pk + lane] = r;
        } else {
            if (lane < 4) {
                const float m = mean[lane];
                const float lm = plog(m);
                bl.c[0][lane] = sqrtf(2.0f * m);
                bl.c[1][lane] = lm;
                bl.c[2][lane] = m * lm - plgamma_approx(m + 1.0f);
            }
            __syncthreads();
            NegBinTangent pol{mean, bl.c[0], bl.c[1], bl.c[2], lg};
            op_reject(e, s.nbt, out + 4 * pk, 4, pol, gd, lane, q);
        }
        __syncthreads();
    }
}

__device__ void run_segment(Eng& e, ProgramLds& s, BetaLds& bl, const Segment& sg, Guard gd,
                            const float* lg, int lane) {
    const int64_t n = sg.count;
    switch (sg.kind) {
        case SEG_U64: {
            RawU64 d{(uint64_t*)sg.out};
            op_fixed(e, s.carry_fixed, n, d, lane);
            break;
        }
        case SEG_UNIFORM_F32: {
            UniformF32 d{(float*)sg.out, (float)sg.d[0], (float)sg.d[1] - (float)sg.d[0], sg.i0};
            op_fixed(e, s.carry_fixed, n, d, lane);
            break;
        }
        case SEG_UNIFORM_F64: {
            UniformF64 d{(double*)sg.out, sg.d[0], sg.d[1] - sg.d[0], sg.i0};
            op_fixed(e, s.carry_fixed, n, d, lane);
            break;
        }
        case SEG_STD_NORMAL_F32: {
            NormalF32<0> d{(float*)sg.out, 0.f, 1.f, nullptr, nullptr};
            op_fixed(e, s.carry_fixed, n, d, lane);
            break;
        }
        case SEG_NORMAL_F32: {
            NormalF32<1> d{(float*)sg.out, (float)sg.d[0], (float)sg.d[1], nullptr, nullptr};
            op_fixed(e, s.carry_fixed, n, d, lane);
            break;
        }
        case SEG_NORMAL_VOP_F32: {
            NormalF32<2> d{(float*)sg.out, 0.f, 1.f, (const float*)sg.in0, (const float*)sg.in1};
            op_fixed(e, s.carry_fixed, n, d, lane);
            break;
        }
        case SEG_EXPONENTIAL_F32: {
            ExponentialF32 d{(float*)sg.out, (float)sg.d[0], 1.0f, 0};
            op_fixed(e, s.carry_fixed, n, d, lane);
            break;
        }
        case SEG_CAUCHY_F32: {
            CauchyF32 d{(float*)sg.out, (float)sg.d[0], (float)sg.d[1]};
            op_fixed(e, s.carry_fixed, n, d, lane);
            break;
        }
        case SEG_STUDENT_T_F32: {
            const float nu = (float)sg.d[0];
            StudentTF32 d{(float*)sg.out, nu, -2 / nu};
            op_fixed(e, s.carry_fixed, n, d, lane);
            break;
        }
        case SEG_GAMMA_F32:
            run_gamma(e, s, (float*)sg.out, n, (float)sg.d[0], (float)sg.d[1], gd, lane);
            break;
        case SEG_POISSON_I32: {
            // d: mean, ne_mean, sqrt_tmean, log_mean, g1
            if (sg.d[0] < 12) {
                op_knuth(e, s.knuth.carry, s.knuth.u, (int32_t*)sg.out, n, (float)sg.d[1], 0, 0, gd, lane);
            } else {
                PoissonTangent pol{(float)sg.d[2], (float)sg.d[0], (float)sg.d[3], (float)sg.d[4], lg};
                op_reject(e, s.pt, (int32_t*)sg.out, n, pol, gd, lane);
            }
            break;
        }
        case SEG_BINOMIAL_I32: {
            // i0: trials; d: p, small_p, mean, g1, sqrt_v, log_small_p, log_small_q, e^-mean
            const int32_t trials = sg.i0;
            const int flip = (sg.d[0] == sg.d[1]) ? 0 : 1;
            if (trials < 25) {
                op_binomial_small(e, s.carry_beta, (int32_t*)sg.out, n, trials, (float)sg.d[0], lane);
            } else if (sg.d[2] < 1.0) {
                op_knuth(e, s.knuth.carry, s.knuth.u, (int32_t*)sg.out, n, (float)sg.d[7], trials, flip, gd, lane);
            } else {
                BinomialTangent pol{(float)trials, (float)sg.d[4], (float)sg.d[2], (float)sg.d[5],
                                    (float)sg.d[6], (float)sg.d[3], trials, flip, lg};
                op_reject(e, s.bt, (int32_t*)sg.out, n, pol, gd, lane);
            }
            break;
        }
        case SEG_BINOMIAL_VOP_I32:
            op_binomial_vop(e, s.vop.carry, (int32_t*)sg.out, n, Param<int32_t>{(const int32_t*)sg.in0, 0},
                            Param<float>{(const float*)sg.in1, 0.f}, gd, lg, lane);
            break;
        case SEG_GAMMA_VOP_F32:
            op_gamma_vop(e, s.vop.carry, (float*)sg.out, n, (const float*)sg.in0,
                         (const float*)sg.in1, gd, lane);
            break;
        case SEG_POISSON_VOP_I32:
            op_poisson_vop(e, s.vop.carry, (int32_t*)sg.out, n, (const float*)sg.in0, gd,
                           lg, lane);
            break;
        case SEG_DISCARD:
            run_discard(e, n, lane);
            break;
        case SEG_DISCRETE_I32:
        case SEG_MULTINOMIAL_SMALL: {
            DiscreteTable t;
            t.ncdf = (int32_t)sg.d[1];
            t.cdf = (const uint32_t*)sg.in0;
            t.prob = (const int32_t*)sg.in0;
            t.alias = (const int32_t*)sg.in1;
            t.bitmask = (uint32_t)sg.d[2];
            if (sg.kind == SEG_DISCRETE_I32) {
                DiscreteUnits d{(int32_t*)sg.out, t};
                op_fixed(e, s.carry_fixed, n, d, lane);
            } else {
                op_multinomial_small(e, (int32_t*)sg.out, n, sg.i0, (int32_t)sg.d[0], t, lane);
            }
            break;
        }
        case SEG_LOGNORMAL_F32: {
            NormalF32<3> d{(float*)sg.out, (float)sg.d[0], (float)sg.d[1], nullptr, nullptr, nullptr};
            op_fixed(e, s.carry_fixed, n, d, lane);
            break;
        }
        case SEG_WEIBULL_F32: {
            WeibullF32 d{(float*)sg.out, (float)sg.d[0], (float)sg.d[1]};
            op_fixed(e, s.carry_fixed, n, d, lane);
            break;
        }
        case SEG_EXTREME_VALUE_F32: {
            ExtremeValueF32 d{(float*)sg.out, (float)sg.d[0], (float)sg.d[1]};
            op_fixed(e, s.carry_fixed, n, d, lane);
            break;
        }
        case SEG_NEGATIVE_BINOMIAL_I32:
            run_negbin(e, s, bl, (int32_t*)sg.out, n, (float)sg.d[0], (float)sg.d[1], gd, lg, lane);
            break;
        case SEG_BETA_F32:
            run_beta(e, s, bl, (float*)sg.out, n, (float)sg.d[0], (float)sg.d[1], (float)sg.d[2], gd, lane);
            break;
        case SEG_MULTINOMIAL_ROWS: {
            // MultinomialGen's large-trials path (Multinomial.h:74-91) for the
            // samples of one engine: row i is BinomialVGen<float>(r, p_i)
            // over the samples (row 0 with every sample's trials), r -= row;
            // the last category takes what remains.
            const int64_t K = (int64_t)sg.d[1];
            const float* prow = (const float*)sg.in1;
            int32_t* rr = (int32_t*)sg.in0;
            int32_t* o = (int32_t*)sg.out;
            // Row constants for 64 rows at a time, one row per thread, read
            // per row with readlane.
            int32_t rreg = sg.i0;  // n <= 64: thread i keeps sample i's remaining trials
            float c_p = 0.f, c_ps = 0.f, c_lp = 0.f, c_lq = 0.f;
            for (int64_t row = 0; row + 1 < K; ++row) {
                const int q = (int)(row & 63);
                if (q == 0) {
                    c_p = row + lane + 1 < K ? prow[row + lane] : 0.5f;
                    c_ps = (c_p < 1.0f - c_p) ? c_p : 1.0f - c_p;  // pmin(p, 1-p)
                    c_lp = plog(c_ps);
                    c_lq = plog(1.0f - c_ps);
                }
                const float rp = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, c_p), q));
                const float rps = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, c_ps), q));
                const float rlp = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, c_lp), q));
                const float rlq = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, c_lq), q));
                const int32_t t0 = row == 0 ? sg.i0 : -1;
                if (n <= 64)
                    op_binomial_row8<true>(e, s.vop.carry, o + row, K, rr, rreg, t0, n, rp, rps, rlp, rlq, gd,
                                           lg, lane);
                else
                    op_binomial_row8<false>(e, s.vop.carry, o + row, K, rr, rreg, t0, n, rp, rps, rlp, rlq,
                                            gd, lg, lane);
            }
            if (n <= 64) {
                if (lane < n) o[lane * K + (K - 1)] = rreg;
            } else {
                for (int64_t i = lane; i < n; i += 64) o[i * K + (K - 1)] = K == 1 ? sg.i0 : rr[i];
            }
            break;
        }
        default:
            if (lane == 0) atomicOr(gd.status, kStatusBadSegment);
            break;
    }
    __syncthreads();
}

// Group b runs segments [begin[b], begin[b+1]) on engine engine_of[b].
__global__ __launch_bounds__(64) void k_program(EngineState* __restrict__ states,
                                                const int32_t* __restrict__ engine_of,
                                                const int32_t* __restrict__ begin,
                                                const Segment* __restrict__ segs,
                                                Guard gd,
                                                const float* __restrict__ lg) {
    __shared__ uint64_t mt[kWords];
    __shared__ ProgramLds s;
    __shared__ BetaLds bl;
    const int lane = threadIdx.x;
    const int b = blockIdx.x;
    EngineState* st = states + engine_of[b];
    Eng e;
    eng_load(e, mt, st, lane);
    for (int k = begin[b]; k < begin[b + 1]; ++k) run_segment(e, s, bl, segs[k], gd, lg, lane);
    eng_store(e, st, lane);
}

// Reload a P8 adaptor's two Vmt states from their records and advance the
// scalar stream by `used` scalars from the records' position (the position
// the feed reached is past that: it fills whole blocks of draws ahead).
__device__ void p8_seek(P8Feed& f, const EngineState* st0, const EngineState* st1, int64_t used, int lane) {
    __syncthreads();
    load_state(f.v0, st0, lane);
    load_state(f.v1, st1, lane);
    int64_t o = p8_offset(st0->idx, (int)st0->reserved) + used;
    while (((o + 7) >> 3) > kN) {
        twist_lds(f.v0, lane);
        twist_lds(f.v1, lane);
        o -= 8 * kN;
    }
    f.idx = (int)((o + 7) >> 3);
    f.b = (int)(o & 7);
}

// Segment programs on P8_mt19937_64 adaptors: group b runs on adaptor
// engine_of[b] = records 2a (Vmt seed) and 2a + 1 (Vmt seed + 4). The
// segment's draws come packed from the adaptor's scalar stream (P8Feed;
// mode >= 0 forces a packing, else the kind's own); afterwards the adaptor
// is left exactly as far along its stream as the scalars consumed. Record 2a
// keeps the operator() buffer position (scalars of the last refill used) in
// its `reserved` word.
__global__ __launch_bounds__(64) void k_program_p8(EngineState* __restrict__ states,
                                                   const int32_t* __restrict__ engine_of,
                                                   const int32_t* __restrict__ begin,
                                                   const Segment* __restrict__ segs, int mode_override,
                                                   Guard gd, const float* __restrict__ lg) {
    __shared__ uint64_t mt[kWords];
    __shared__ uint64_t vst[2][kWords];
    __shared__ ProgramLds s;
    __shared__ BetaLds bl;
    const int lane = threadIdx.x;
    const int b = blockIdx.x;
    EngineState* st0 = states + 2 * (int64_t)engine_of[b];
    EngineState* st1 = st0 + 1;
    P8Feed f;
    f.v0 = &vst[0][0];
    f.v1 = &vst[1][0];
    f.mode = kP8Raw;
    f.refills = 0;
    int64_t used = 0;  // scalars consumed by the segments so far
    p8_seek(f, st0, st1, 0, lane);
    __syncthreads();
    for (int k = begin[b]; k < begin[b + 1]; ++k) {
        const int mode = mode_override >= 0 ? mode_override : p8_feed_mode(segs[k].kind);
        if (mode < 0) {
            if (lane == 0) atomicOr(gd.status, kStatusBadSegment);
            continue;
        }
        Eng e;
        e.mt = mt;
        e.idx = kN;  // the first draw refills the block from the feed
        e.half_valid = 0;
        e.half_cache[0] = e.half_cache[1] = 0;
        e.twisted = false;
        e.feed = &f;
        f.mode = mode;
        f.refills = 0;
        run_segment(e, s, bl, segs[k], gd, lg, lane);
        // slots of the last packed block used: full draws, or half draws
        // (an odd count leaves the last draw's upper half as the cache)
        const int64_t last = mode == kP8H2 ? 2 * (2 * (int64_t)e.idx - e.half_valid)
                           : mode == kP8F34 ? 3 * (int64_t)e.idx : 4 * (int64_t)e.idx;
        if (f.refills > 0) used += (int64_t)(f.refills - 1) * p8_block_scalars(mode) + last;
        p8_seek(f, st0, st1, used, lane);  // the next segment continues from there
        __syncthreads();
    }
    store_state(st0, f.v0, lane);
    store_state(st1, f.v1, lane);
    if (lane == 0) {
        st0->idx = st1->idx = f.idx;
        st0->reserved = (uint64_t)f.b;
    }
}

hipError_t launch_program_p8(EngineState* st, int32_t n_groups, const int32_t* d_engine_of, const int32_t* d_begin,
                             const Segment* d_segs, int mode_override, Guard gd, hipStream_t s) {
    if (n_groups <= 0) return hipSuccess;
    const float* lg = lgamma_table(s);
    if (!lg) return hipErrorOutOfMemory;
    hipLaunchKernelGGL(k_program_p8, dim3((unsigned)n_groups), dim3(64), 0, s, st, d_engine_of, d_begin, d_segs,
                       mode_override, gd, lg);
    return hipGetLastError();
}

hipError_t launch_program(EngineState* st, int32_t n_groups, const int32_t* d_engine_of,
                          const int32_t* d_begin, const Segment* d_segs, Guard gd,
                          hipStream_t s) {
    if (n_groups <= 0) return hipSuccess;
    const float* lg = lgamma_table(s);
    if (!lg) return hipErrorOutOfMemory;
    hipLaunchKernelGGL(k_program, dim3((unsigned)n_groups), dim3(64), 0, s, st, d_engine_of, d_begin,
                       d_segs, gd, lg);
    return hipGetLastError();
}

// ------------------------------------------------------ host: segments ---
// Constructor constants of the reference generators, computed on the host in
// double exactly as the reference constructors do.

Segment seg_poisson(int32_t* out, int64_t count, double mean) {
    Segment s{};
    s.kind = SEG_POISSON_I32;
    s.count = count;
    s.out = out;
    s.d[0] = mean;                                   // PoissonGen (Discrete.h:807-813)
    s.d[1] = std::exp(-mean);
    s.d[2] = std::sqrt(2 * mean);
    s.d[3] = std::log(mean);
    s.d[4] = mean * s.d[3] - std::lgamma(mean + 1);
    return s;
}

Segment seg_binomial(int32_t* out, int64_t count, int32_t trials, double p) {
    Segment s{};
    s.kind = SEG_BINOMIAL_I32;
    s.count = count;
    s.out = out;
    s.i0 = trials;                                   // BinomialGen (Discrete.h:929-941)
    s.d[0] = p;
    s.d[1] = std::min(p, 1 - p);
    s.d[2] = trials * s.d[1];
    s.d[7] = std::exp(-s.d[2]);                      // PoissonGen{mean}.ne_mean
    if (!(trials < 25 || s.d[2] < 1.0)) {
        s.d[3] = std::lgamma(trials + 1);
        s.d[4] = std::sqrt(2 * s.d[2] * (1 - s.d[1]));
        s.d[5] = std::log(s.d[1]);
        s.d[6] = std::log(1 - s.d[1]);
    }
    return s;
}

}  // namespace erd
