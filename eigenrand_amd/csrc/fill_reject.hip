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
#include "erd_device.hpp"
#include "erd_kernels.hpp"
#include "erd_slots.hpp"

namespace erd {

constexpr int kRejectCap = 1 << 22;  // candidates without output before giving up

// ------------------------------------------------------------- policies ---
// W: lanes per packet; SW: words per slot (4 full draw, 2 half draw);
// K: slots per candidate; eval(w, h, v, acc): lane 2p+h of a candidate from
// word p of each of its K slots; store(out, i, v): write output element i.

struct GammaGS {  // NormalExp.h:536-558
    static constexpr int W = 8, SW = 4, K = 3;
    using Out = float;
    float beta, px, inv_alpha, am1;
    __device__ void eval(const uint64_t* w, int h, float& v, bool& acc) const {
        const float ux = ur_f32(h ? hi32(w[0]) : lo32(w[0]));
        const float vx = nzur_f32(h ? hi32(w[1]) : lo32(w[1]));
        const float lv = plog_pos(vx);
        const float x1 = pexp(inv_alpha * lv);
        const float q1 = pexp(-x1);
        const float x2 = 1.0f - lv;
        const float q2 = pexp(plog_pos(x2) * am1);
        const bool c = ux < px;
        const float xx = c ? x1 : x2;
        const float qx = c ? q1 : q2;
        const float u3 = ur_f32(h ? hi32(w[2]) : lo32(w[2]));
        v = beta * xx;
        acc = u3 < qx;
    }
    __device__ void store(Out* o, int64_t i, float v) const { o[i] = v; }
};

struct GammaTangent {  // NormalExp.h:580-603
    static constexpr int W = 8, SW = 4, K = 2;
    using Out = float;
    float beta, sq, am1;
    __device__ void eval(const uint64_t* w, int h, float& v, bool& acc) const {
        float ys, yc;
        psincos(ERD_F(kPi) * ur_f32(h ? hi32(w[0]) : lo32(w[0])), ys, yc);
        const float yx = ys / yc;
        const float xx = sq * yx + am1;
        const bool c = 0.0f < xx;
        const float ux = ur_f32(h ? hi32(w[1]) : lo32(w[1]));
        const float ub = (yx * yx + 1.0f) * pexp(am1 * plog(xx / am1) - yx * sq);
        acc = c && (ux <= ub);
        v = beta * xx;
    }
    __device__ void store(Out* o, int64_t i, float v) const { o[i] = v; }
};

struct PoissonTangent {  // Discrete.h:875-902 (Packet4f via half())
    static constexpr int W = 4, SW = 2, K = 2;
    using Out = int32_t;
    float psq, pmean, plm, pg1;
    __device__ void eval(const uint64_t* w, int h, float& v, bool& acc) const {
        float s, c;
        psincos(ERD_F(kPi) * ur_f32(h ? hi32(w[0]) : lo32(w[0])), s, c);
        const float yx = s / c;
        const float fres = truncf(psq * yx + pmean);
        const float p1 = (yx * yx + 1.0f) * ERD_F(0.9);
        const float p2 = pexp((fres * plm - plgamma_approx(fres + 1.0f)) - pg1);
        const bool c1 = 0.0f <= fres;
        const float u2 = ur_f32(h ? hi32(w[1]) : lo32(w[1]));
        acc = c1 && (u2 <= p1 * p2);
        v = fres;
    }
    __device__ void store(Out* o, int64_t i, float v) const { o[i] = (int32_t)v; }
};

struct BinomialTangent {  // Discrete.h:1011-1051
    static constexpr int W = 4, SW = 2, K = 2;
    using Out = int32_t;
    float ptrials, psqrt_v, pmean, plp, plq, pg1;
    int32_t trials;
    int flip;  // p != small_p: result is trials - k
    __device__ void eval(const uint64_t* w, int h, float& v, bool& acc) const {
        float s, c;
        psincos(ERD_F(kPi) * ur_f32(h ? hi32(w[0]) : lo32(w[0])), s, c);
        const float ys = s / c;
        const float fres = truncf(psqrt_v * ys + pmean);
        const float p1 = (ERD_F(1.2) * psqrt_v) * (1.0f + ys * ys);
        const float p2 = pexp((((pg1 - plgamma_approx(fres + 1.0f)) -
                                plgamma_approx((ptrials + 1.0f) - fres)) +
                               fres * plp) +
                              (ptrials - fres) * plq);
        const bool c1 = (0.0f <= fres) && (fres <= ptrials);
        const float u2 = ur_f32(h ? hi32(w[1]) : lo32(w[1]));
        acc = c1 && (u2 <= p1 * p2);
        v = fres;
    }
    __device__ void store(Out* o, int64_t i, float v) const {
        const int32_t k = (int32_t)v;
        o[i] = flip ? trials - k : k;
    }
};

// --------------------------------------------------------------- kernel ---
template <class P>
__global__ __launch_bounds__(64) void k_fill_reject(EngineState* __restrict__ states,
                                                    typename P::Out* __restrict__ out, int64_t per,
                                                    P pol, uint32_t* __restrict__ status) {
    constexpr int W = P::W, SW = P::SW, K = P::K;
    constexpr int IPC = W / 2;          // threads per candidate (2 lanes each)
    constexpr int CPS = 64 / IPC;       // candidates per wave step
    constexpr int SLOTS = SlotMode<SW>::kSlots;
    // Carried slots: at most K-1 (a straddling candidate's first slots, or the
    // engine's cached half, which the first candidate always consumes).
    constexpr int MAXC = K - 1;
    __shared__ uint64_t mt[kWords];
    __shared__ uint64_t carry[MAXC * SW];
    __shared__ float rest[W];
    __shared__ int32_t tab[CPS + 1];

    const int lane = threadIdx.x;
    const int64_t g = blockIdx.x;
    EngineState* st = states + g;
    typename P::Out* o = out + g * per;
    const int64_t T = per / W;  // packets to emit
    if (T == 0) return;

    load_state(mt, st, lane);
    int nc = 0;
    int pos;
    if (SW == 2) {
        pos = 2 * st->idx;
        if (st->half_valid) {
            if (lane < 2) carry[lane] = st->half_cache[lane];
            nc = 1;
        }
        __syncthreads();
    } else {
        pos = st->idx;
    }
    bool twisted = false;

    int64_t full_cnt = 0;  // candidates emitted whole
    int64_t fifo = 0;      // queue items enqueued so far
    int64_t since_emit = 0;
    int end_v = -1;        // virtual slot just past the last used candidate
    const int cl = lane / IPC, p = lane % IPC;

    for (;;) {
        if (pos >= SLOTS) {
            twist_lds(mt, lane);
            pos = 0;
            twisted = true;
        }
        const int avail = nc + (SLOTS - pos);
        const int ncand = avail / K;
        int c0 = 0;
        for (; c0 < ncand; c0 += CPS) {
            const int c = c0 + cl;
            const bool live = c < ncand;
            float v[2] = {0.f, 0.f};
            bool a[2] = {false, false};
            if (live) {
                uint64_t w[K];
#pragma unroll
                for (int k = 0; k < K; ++k) w[k] = slot_word<SW>(mt, carry, nc, pos, c * K + k, p);
#pragma unroll
                for (int h = 0; h < 2; ++h) pol.eval(w, h, v[h], a[h]);
            }
            const uint64_t B0 = __ballot(a[0]);
            const uint64_t B1 = __ballot(a[1]);
            const uint64_t cmask = (IPC == 4 ? 0xFull : 0x3ull) << (cl * IPC);
            const bool full = live && ((B0 & B1 & cmask) == cmask);
            const bool q0 = a[0] && !full, q1 = a[1] && !full;
            const uint64_t A0 = __ballot(q0), A1 = __ballot(q1);
            const uint64_t FC = __ballot(full && p == 0);
            const uint64_t below_c = lanemask_lt(cl * IPC);
            const int64_t fb = full_cnt + __popcll(FC & below_c);          // bypasses before c
            const int64_t fstart = fifo + __popcll(A0 & below_c) + __popcll(A1 & below_c);
            const int64_t f0 = fifo + __popcll(A0 & lanemask_lt(lane)) + __popcll(A1 & lanemask_lt(lane));
            const int64_t f1 = f0 + (q0 ? 1 : 0);
            const int64_t fifo_end = fifo + __popcll(A0) + __popcll(A1);
            const int64_t m0 = fifo / W;                      // first queued packet open at step start
            const int64_t m_end = fifo_end / W;               // queued packets complete after the step
            const int acnt = __popcll(A0 & cmask) + __popcll(A1 & cmask);
            // Output packet emitted by this candidate (-1 if none).
            int64_t emit = -1;
            if (live && full) emit = fb + fstart / W;
            else if (live && acnt > 0 && (fstart + acnt) / W > fstart / W) emit = fb + fstart / W;
            // Table of output packets for queued packets completed in this step.
            if (live && !full && emit >= 0 && p == 0) tab[fstart / W - m0] = (int32_t)emit;
            // Candidate that emits the last packet ends the fill.
            const uint64_t LAST = __ballot(emit == T - 1 && p == 0);
            __syncthreads();
            // Bypassing candidates: write the packet directly.
            if (full && emit < T) {
#pragma unroll
                for (int h = 0; h < 2; ++h) pol.store(o, emit * W + 2 * p + h, v[h]);
            }
            // Earlier waiting lanes of queued packet m0, if it completed now.
            const int pend = (int)(fifo % W);
            if (m_end > m0 && lane < pend) {
                const int64_t e = tab[0];
                if (e < T) pol.store(o, e * W + lane, rest[lane]);
            }
            // This step's queued lanes: written out if their packet is complete,
            // otherwise parked in the rest queue.
            float park[2];
            int park_at[2] = {-1, -1};
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const bool q = h ? q1 : q0;
                const int64_t f = h ? f1 : f0;
                if (!q) continue;
                const int64_t m = f / W;
                if (m < m_end) {
                    const int64_t e = tab[m - m0];
                    if (e < T) pol.store(o, e * W + (f % W), v[h]);
                } else {
                    park[h] = v[h];
                    park_at[h] = (int)(f % W);
                }
            }
            __syncthreads();
#pragma unroll
            for (int h = 0; h < 2; ++h)
                if (park_at[h] >= 0) rest[park_at[h]] = park[h];
            full_cnt += __popcll(FC);
            fifo = fifo_end;
            if (LAST) {
                const int clast = (int)(__ffsll((long long)LAST) - 1) / IPC;
                end_v = (c0 + clast + 1) * K;
                break;
            }
            if (__popcll(FC) + (m_end - m0) > 0) {
                since_emit = 0;
            } else {
                since_emit += CPS;
                if (since_emit > kRejectCap) break;
            }
        }
        if (end_v >= 0) break;
        if (since_emit > kRejectCap) {
            if (lane == 0) atomicOr(status, kStatusIterCap);
            end_v = ncand * K;  // leave the engine after the evaluated candidates
            break;
        }
        // Block exhausted: carry the slots of the straddling candidate.
        carry_tail<SW, MAXC>(mt, carry, nc, pos, ncand * K, avail, lane);
        nc = avail - ncand * K;
        pos = SLOTS;
    }
    __syncthreads();
    // Final engine position: virtual slot end_v.
    int bpos;          // block slot position after the fill
    int keep_half = 0;  // half mode: an unconsumed carried half remains
    if (end_v >= nc) {
        bpos = pos + end_v - nc;
    } else {
        bpos = pos;
        keep_half = 1;  // only possible for the engine's own cached half
    }
    if (twisted) store_state(st, mt, lane);
    if (lane == 0) {
        if (SW == 4) {
            st->idx = bpos;
        } else if (keep_half) {
            st->idx = bpos / 2;
        } else if (bpos & 1) {
            st->idx = (bpos + 1) / 2;
            st->half_valid = 1;
            st->half_cache[0] = temper(mt[2 * bpos]);
            st->half_cache[1] = temper(mt[2 * bpos + 1]);
        } else {
            st->idx = bpos / 2;
            st->half_valid = 0;
            // The reference keeps the last split draw's high half in its cache
            // member after handing it out; mirror it so records compare equal.
            if (twisted || end_v > nc) {
                st->half_cache[0] = temper(mt[2 * (bpos - 1)]);
                st->half_cache[1] = temper(mt[2 * (bpos - 1) + 1]);
            }
        }
    }
}

// ------------------------------------------------------------- launchers ---
template <class P>
static hipError_t launch_reject(EngineState* st, int64_t n, typename P::Out* out, int64_t per,
                                const P& pol, uint32_t* status, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_reject<P>, dim3((unsigned)n), dim3(64), 0, s, st, out, per, pol, status);
    return hipGetLastError();
}

hipError_t launch_fill_gamma_reject_f32(EngineState* st, int64_t n, float* out, int64_t per,
                                        float alpha, float beta, uint32_t* status, hipStream_t s) {
    if (alpha < 1.0f) {
        GammaGS pol{beta, (float)(kE / (alpha + kE)), 1 / alpha, alpha - 1};
        return launch_reject(st, n, out, per, pol, status, s);
    }
    GammaTangent pol{beta, sqrtf(2 * alpha - 1), alpha - 1};
    return launch_reject(st, n, out, per, pol, status, s);
}

hipError_t launch_fill_poisson_reject_i32(EngineState* st, int64_t n, int32_t* out, int64_t per,
                                          const PoissonParams& pp, uint32_t* status, hipStream_t s) {
    PoissonTangent pol{(float)pp.sqrt_tmean, (float)pp.mean, (float)pp.log_mean, (float)pp.g1};
    return launch_reject(st, n, out, per, pol, status, s);
}

hipError_t launch_fill_binomial_reject_i32(EngineState* st, int64_t n, int32_t* out, int64_t per,
                                           const BinomialParams& bp, uint32_t* status,
                                           hipStream_t s) {
    BinomialTangent pol{(float)bp.trials, (float)bp.sqrt_v, (float)bp.mean, (float)bp.log_small_p,
                        (float)bp.log_small_q, (float)bp.g1, bp.trials, bp.p == bp.small_p ? 0 : 1};
    return launch_reject(st, n, out, per, pol, status, s);
}

}  // namespace erd
