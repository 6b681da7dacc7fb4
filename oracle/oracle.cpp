// Oracle (test infrastructure only): a plain C++ restatement of EigenRand's
// packet algorithms for the sampling hot path, lane-emulated at the canonical
// AVX2 layout (Packet8f = 8 float lanes fed by one Vmt19937_64 call of 4 u64
// lanes; int32 distributions use Packet4i/Packet4f fed by half-calls).
// Built with -ffp-contract=off so every float op rounds exactly as the
// reference's unfused packet ops do.
//
// Only tests/, the smoke check and the CPU baseline leg of the benchmark may
// load this library; it is never on the product path.
#include <cstdint>
#include <cstring>
#include <cmath>
#include <vector>
#include <thread>
#include <algorithm>
#include <memory>
#include <limits>
#include <random>
#include "orc_math.h"

using namespace orc;

namespace {

// ---------------------------------------------------------------- engine ---
// MT19937-64 constants: EigenRand/PacketRandomEngine.h:421-426 (Pmt19937_64).
constexpr int kN = 312, kM = 156;
constexpr uint64_t kA = 0xb5026f5aa96619e9ull;
constexpr uint64_t kUM = 0xFFFFFFFF80000000ull;  // (~0 << 31)
constexpr uint64_t kLM = 0x7FFFFFFFull;
constexpr uint64_t kF = 6364136223846793005ull;

// Vmt19937_64 under AVX2 = Pmt19937_64<Packet8i>: 4 u64 lanes, lane l is an
// independent mt19937_64 seeded x0+l (PacketRandomEngine.h:230-265).
// Layout shared with the device engine state: word i of lane l at mt[i*4+l].
struct Vmt {
    uint64_t mt[kN * 4];
    int32_t idx;          // next word; kN means "twist before the next draw"
    int32_t half_valid;   // MersenneTwister::half() cache (PacketRandomEngine.h:325-336)
    uint64_t half_cache[2];
    uint64_t reserved;    // pads the record to a multiple of 16 bytes
};
static_assert(sizeof(Vmt) == 10016, "engine state layout");

void vmt_seed(Vmt& e, uint64_t x0) {
    // seed(): state[i] = f*(prev ^ prev>>62) + i (PacketRandomEngine.h:256-265,
    // pmuluadd64 arch/AVX/MorePacketMath.h:684-694).
    for (int l = 0; l < 4; ++l) {
        uint64_t prev = x0 + (uint64_t)l;
        e.mt[l] = prev;
        for (int i = 1; i < kN; ++i) {
            prev = (prev ^ (prev >> 62)) * kF + (uint64_t)i;
            e.mt[i * 4 + l] = prev;
        }
    }
    e.idx = kN;
    e.half_valid = 0;
    e.half_cache[0] = e.half_cache[1] = 0;
    e.reserved = 0;
}

void vmt_twist(Vmt& e) {
    // refill_upper/refill_lower (PacketRandomEngine.h:340-404) write the next
    // 312 words of each lane; in single-buffer form that is the standard twist
    // x[i] = x[i+m] ^ twist(x[i], x[i+1]), indices mod n, in increasing i.
    // Written as the usual three loops (no modulo) with the 4 lanes innermost.
    uint64_t* x = e.mt;
    auto tw = [](uint64_t cur, uint64_t nxt, uint64_t far) {
        uint64_t y = (cur & kUM) | (nxt & kLM);
        return far ^ (y >> 1) ^ ((0 - (y & 1)) & kA);
    };
    int i = 0;
    for (; i < kN - kM; ++i)
        for (int l = 0; l < 4; ++l) x[i * 4 + l] = tw(x[i * 4 + l], x[(i + 1) * 4 + l], x[(i + kM) * 4 + l]);
    for (; i < kN - 1; ++i)
        for (int l = 0; l < 4; ++l) x[i * 4 + l] = tw(x[i * 4 + l], x[(i + 1) * 4 + l], x[(i + kM - kN) * 4 + l]);
    for (int l = 0; l < 4; ++l) x[i * 4 + l] = tw(x[i * 4 + l], x[l], x[(kM - 1) * 4 + l]);
}

inline uint64_t temper(uint64_t y) {
    // operator() tempering: PacketRandomEngine.h:303-308
    y ^= (y >> 29) & 0x5555555555555555ull;
    y ^= (y << 17) & 0x71d67fffeda60000ull;
    y ^= (y << 37) & 0xfff7eee000000000ull;
    y ^= y >> 43;
    return y;
}

// A record with half_valid < 0 stands for a scalar engine fed into packet
// code (P8_mt19937_64 through RawbitsMaker, below); `reserved` points to it.
struct P8Scalar;
void p8s_full(P8Scalar* f, uint64_t out[4]);
void p8s_half(P8Scalar* f, uint64_t out[2]);

inline void vmt_draw(Vmt& e, uint64_t out[4]) {
    if (e.half_valid < 0) { p8s_full(reinterpret_cast<P8Scalar*>(e.reserved), out); return; }
    if (e.idx >= kN) { vmt_twist(e); e.idx = 0; }
    for (int l = 0; l < 4; ++l) out[l] = temper(e.mt[e.idx * 4 + l]);
    e.idx++;
}

// half(): low 128 bits now, high 128 bits cached (PacketRandomEngine.h:325-336).
inline void vmt_half(Vmt& e, uint64_t out[2]) {
    if (e.half_valid < 0) { p8s_half(reinterpret_cast<P8Scalar*>(e.reserved), out); return; }
    if (e.half_valid) {
        e.half_valid = 0;
        out[0] = e.half_cache[0]; out[1] = e.half_cache[1];
        return;
    }
    uint64_t p[4];
    vmt_draw(e, p);
    out[0] = p[0]; out[1] = p[1];
    e.half_cache[0] = p[2]; e.half_cache[1] = p[3];
    e.half_valid = 1;
}

// Packet8f uniform_real: rawbits_34 = one full call (arch/AVX/RandUtils.h:
// 154-157, 206-209); float j = 32-bit word j = u64 lane j/2, low word first.
inline void ur8(Vmt& e, float u[8]) {
    uint64_t p[4];
    vmt_draw(e, p);
    for (int j = 0; j < 8; ++j) u[j] = bit_to_ur_float((uint32_t)(p[j >> 1] >> (32 * (j & 1))));
}
inline void nzur8(Vmt& e, float u[8]) {
    ur8(e, u);
    for (int j = 0; j < 8; ++j) u[j] = u[j] + 1.1920928955078125e-07f / 8;
}
// Packet4f uniform_real via half() (arch/AVX/RandUtils.h:21-38).
inline void ur4(Vmt& e, float u[4]) {
    uint64_t p[2];
    vmt_half(e, p);
    for (int j = 0; j < 4; ++j) u[j] = bit_to_ur_float((uint32_t)(p[j >> 1] >> (32 * (j & 1))));
}

// ---------------------------------------------------------- CompressMask ---
// Lane-compaction FIFO of arch/AVX/PacketFilter.h:147-184 (W=8) and
// arch/SSE/PacketFilter.h:118-150 (W=4): a fully accepted packet is emitted as
// is and leaves the queue untouched; otherwise accepted lanes are appended in
// lane order and the first W queued values are emitted once W are available.
template <int W, class T = float>
bool compress_append(T value[W], const bool mask[W], T rest[W], int& rest_cnt) {
    int cnt = 0;
    for (int i = 0; i < W; ++i) cnt += mask[i];
    if (cnt == W) return true;
    T q[2 * W];
    int n = 0;
    for (int i = 0; i < rest_cnt; ++i) q[n++] = rest[i];
    for (int i = 0; i < W; ++i) if (mask[i]) q[n++] = value[i];
    if (n >= W) {
        for (int i = 0; i < W; ++i) value[i] = q[i];
        for (int i = W; i < n; ++i) rest[i - W] = q[i];
        rest_cnt = n - W;
        return true;
    }
    for (int i = 0; i < n; ++i) rest[i] = q[i];
    rest_cnt = n;
    return false;
}

constexpr float kPiF = (float)kPi;

// ------------------------------------------------------------ generators ---
// StdNormalGen::packetOp (EigenRand/Dists/NormalExp.h:60-82): Box-Muller with
// the cos half cached for the next packet.
struct StdNormal8 {
    bool valid = false;
    float cache[8];
    void op(Vmt& e, float out[8]) {
        if (valid) { valid = false; std::memcpy(out, cache, sizeof(cache)); return; }
        valid = true;
        float u1[8], u2[8];
        ur8(e, u1);
        ur8(e, u2);
        for (int j = 0; j < 8; ++j) {
            float a = 1.0f - u1[j];
            float radius = std::sqrt(-2.0f * plog(a));
            float theta = F(2 * kPi) * u2[j];
            float s, c;
            psincos(theta, s, c);
            cache[j] = radius * c;
            out[j] = radius * s;
        }
    }
};

// GammaGen::packetOp (NormalExp.h:527-605); ctor constants 453-458.
template <int W>
inline void urW(Vmt& e, float* u);

// GammaGen<float>::packetOp at W lanes: Packet8f (W = 8, full draws) or
// Packet4f (W = 4, half draws through half(), as NegativeBinomialGen uses it);
// the CompressMask queue is `rest`.
template <int W>
inline void nzurW(Vmt& e, float* u) {
    if (W == 4) {
        ur4(e, u);
        for (int j = 0; j < 4; ++j) u[j] = u[j] + 1.1920928955078125e-07f / 8;
    } else {
        nzur8(e, u);
    }
}

template <int W>
struct GammaPk {
    float alpha, beta, px, sq;
    float rest[W];
    int rest_cnt = 0;
    GammaPk(float a, float b) : alpha(a), beta(b) {
        px = (float)(kE / (alpha + kE));
        sq = std::sqrt(2 * alpha - 1);
    }
    void op(Vmt& e, float out[W]) {
        if (alpha < 1) {
            for (;;) {
                float ux[W], vx[W], u3[W], cands[W];
                bool acc[W];
                urW<W>(e, ux);
                nzurW<W>(e, vx);
                float xx[W], qx[W];
                for (int j = 0; j < W; ++j) {
                    float lv = plog(vx[j]);
                    float x1 = pexp((1 / alpha) * lv);
                    float q1 = pexp(-x1);
                    float x2 = 1.0f - plog(vx[j]);
                    float q2 = pexp(plog(x2) * (alpha - 1));
                    bool c = ux[j] < px;
                    xx[j] = c ? x1 : x2;
                    qx[j] = c ? q1 : q2;
                }
                urW<W>(e, u3);
                for (int j = 0; j < W; ++j) { cands[j] = beta * xx[j]; acc[j] = u3[j] < qx[j]; }
                if (compress_append<W>(cands, acc, rest, rest_cnt)) { std::memcpy(out, cands, 4 * W); return; }
            }
        }
        if (alpha == 1) {
            float u[W];
            urW<W>(e, u);
            for (int j = 0; j < W; ++j) out[j] = beta * -(plog(1.0f - u[j]) / 1.0f);
            return;
        }
        int count;
        if ((count = (int)alpha) == alpha && count < 20) {
            float yx[W], u[W];
            nzurW<W>(e, yx);
            while (--count) {
                nzurW<W>(e, u);
                for (int j = 0; j < W; ++j) yx[j] = yx[j] * u[j];
            }
            for (int j = 0; j < W; ++j) out[j] = -(beta * plog(yx[j]));
            return;
        }
        for (;;) {
            const float alpha_1 = alpha - 1;
            float u1[W], u2[W], cands[W];
            bool acc[W];
            urW<W>(e, u1);
            float yx[W], xx[W];
            bool c[W];
            for (int j = 0; j < W; ++j) {
                float ys, yc;
                psincos(kPiF * u1[j], ys, yc);
                yx[j] = ys / yc;
                xx[j] = sq * yx[j] + alpha_1;
                c[j] = 0.0f < xx[j];
            }
            urW<W>(e, u2);
            for (int j = 0; j < W; ++j) {
                float ub = (yx[j] * yx[j] + 1.0f) *
                           pexp(alpha_1 * plog(xx[j] / alpha_1) - yx[j] * sq);
                acc[j] = c[j] && (u2[j] <= ub);
                cands[j] = beta * xx[j];
            }
            if (compress_append<W>(cands, acc, rest, rest_cnt)) { std::memcpy(out, cands, 4 * W); return; }
        }
    }
};
using Gamma8 = GammaPk<8>;

// PoissonGen (EigenRand/Dists/Discrete.h:788-904) at Packet4i/Packet4f.
struct Poisson4 {
    double mean, ne_mean, sqrt_tmean, log_mean, g1;
    float rest[4];
    int rest_cnt = 0;
    explicit Poisson4(double m) : mean(m), ne_mean(std::exp(-m)) {
        sqrt_tmean = std::sqrt(2 * mean);
        log_mean = std::log(mean);
        g1 = mean * log_mean - std::lgamma(mean + 1);
    }
    void op(Vmt& e, int32_t out[4]) {
        if (mean < 12) {
            int32_t res[4] = {0, 0, 0, 0};
            float val[4] = {1, 1, 1, 1};
            const float thr = (float)ne_mean;
            for (;;) {
                float u[4];
                ur4(e, u);
                bool any = false;
                bool c[4];
                for (int j = 0; j < 4; ++j) { val[j] = val[j] * u[j]; c[j] = thr < val[j]; any |= c[j]; }
                if (!any) break;
                for (int j = 0; j < 4; ++j) res[j] += c[j] ? 1 : 0;
            }
            std::memcpy(out, res, 16);
            return;
        }
        const float psq = (float)sqrt_tmean, pmean = (float)mean, plm = (float)log_mean, pg1 = (float)g1;
        for (;;) {
            float u1[4], u2[4], fres[4], p12[4];
            bool c1[4], acc[4];
            ur4(e, u1);
            for (int j = 0; j < 4; ++j) {
                float s, c;
                psincos(kPiF * u1[j], s, c);
                float yx = s / c;
                fres[j] = ptruncate(psq * yx + pmean);
                float p1 = (yx * yx + 1.0f) * F(0.9);
                float p2 = pexp((fres[j] * plm - plgamma_approx(fres[j] + 1.0f)) - pg1);
                c1[j] = 0.0f <= fres[j];
                p12[j] = p1 * p2;
            }
            ur4(e, u2);
            for (int j = 0; j < 4; ++j) acc[j] = c1[j] && (u2[j] <= p12[j]);
            if (compress_append<4>(fres, acc, rest, rest_cnt)) {
                for (int j = 0; j < 4; ++j) out[j] = (int32_t)fres[j];
                return;
            }
        }
    }
};

// BinomialGen (Discrete.h:911-1053) at Packet4i/Packet4f; the rejection queue
// is the one owned by its PoissonGen member.
struct Binomial4 {
    Poisson4 poisson;
    int32_t trials;
    double p, small_p, g1 = 0, sqrt_v = 0, log_small_p = 0, log_small_q = 0;
    Binomial4(int32_t n, double pp)
        : poisson(n * std::min(pp, 1 - pp)), trials(n), p(pp), small_p(std::min(pp, 1 - pp)) {
        if (!(trials < 25 || poisson.mean < 1.0)) {
            g1 = std::lgamma(trials + 1);
            sqrt_v = std::sqrt(2 * poisson.mean * (1 - small_p));
            log_small_p = std::log(small_p);
            log_small_q = std::log(1 - small_p);
        }
    }
    void op(Vmt& e, int32_t out[4]) {
        int32_t res[4];
        if (trials < 25) {
            const float pp = (float)p;
            for (int j = 0; j < 4; ++j) res[j] = trials;
            for (int i = 0; i < trials; ++i) {
                float u[4];
                ur4(e, u);
                for (int j = 0; j < 4; ++j) res[j] += (pp <= u[j]) ? -1 : 0;
            }
            std::memcpy(out, res, 16);
            return;
        } else if (poisson.mean < 1.0) {
            poisson.op(e, res);
        } else {
            const float ptrials = (float)trials, psqrt_v = (float)sqrt_v, pmean = (float)poisson.mean,
                        plp = (float)log_small_p, plq = (float)log_small_q, pg1 = (float)g1;
            for (;;) {
                float u1[4], u2[4], fres[4], p12[4];
                bool c1[4], acc[4];
                ur4(e, u1);
                for (int j = 0; j < 4; ++j) {
                    float s, c;
                    psincos(kPiF * u1[j], s, c);
                    float ys = s / c;
                    fres[j] = ptruncate(psqrt_v * ys + pmean);
                    float p1 = (F(1.2) * psqrt_v) * (1.0f + ys * ys);
                    float p2 = pexp((((pg1 - plgamma_approx(fres[j] + 1.0f)) -
                                      plgamma_approx((ptrials + 1.0f) - fres[j])) +
                                     fres[j] * plp) +
                                    (ptrials - fres[j]) * plq);
                    c1[j] = (0.0f <= fres[j]) && (fres[j] <= ptrials);
                    p12[j] = p1 * p2;
                }
                ur4(e, u2);
                for (int j = 0; j < 4; ++j) acc[j] = c1[j] && (u2[j] <= p12[j]);
                if (compress_append<4>(fres, acc, poisson.rest, poisson.rest_cnt)) {
                    for (int j = 0; j < 4; ++j) res[j] = (int32_t)fres[j];
                    break;
                }
            }
        }
        for (int j = 0; j < 4; ++j) out[j] = (p == small_p) ? res[j] : trials - res[j];
    }
};

// BinomialVGen<int32_t>::packetOp (Discrete.h:1124-1192): per-lane float
// parameters, lockstep Knuth loop for lanes with mean < 1, then lockstep
// tangent rejection keeping each lane's first accepted candidate.
template <int W>
inline void urW(Vmt& e, float* u) {
    if (W == 4) ur4(e, u); else ur8(e, u);
}

// W = 4: BinomialVGen<int32_t> at Packet4i/Packet4f (half draws); W = 8:
// BinomialVGen<float> at Packet8f (full draws, Discrete.h:1251-1330, used by
// MultinomialGen with int32 trials stored in float lanes).
template <int W>
void binomialv(Vmt& e, const int32_t trials[W], const float p[W], int32_t out[W]) {
    float psmall[W], ptr[W], pmean[W], pg1[W], psq[W], plp[W], plq[W];
    bool valid[W];
    float valid_res[W] = {0};
    for (int j = 0; j < W; ++j) {
        psmall[j] = (p[j] < 1.0f - p[j]) ? p[j] : 1.0f - p[j];   // pmin(p, 1-p)
        ptr[j] = (float)trials[j];
        pmean[j] = ptr[j] * psmall[j];
        pg1[j] = plgamma_approx(ptr[j] + 1.0f);
        psq[j] = std::sqrt((2.0f * pmean[j]) * (1.0f - psmall[j]));
        plp[j] = plog(psmall[j]);
        plq[j] = plog(1.0f - psmall[j]);
        valid[j] = pmean[j] < 1.0f;
    }
    bool any_valid = false, all_valid = true;
    for (int j = 0; j < W; ++j) { any_valid |= valid[j]; all_valid &= valid[j]; }
    if (any_valid) {
        int32_t res[W] = {0};
        float val[W], nem[W];
        for (int j = 0; j < W; ++j) val[j] = 1.0f;
        for (int j = 0; j < W; ++j) nem[j] = pexp(-pmean[j]);
        for (;;) {
            float u[W];
            urW<W>(e, u);
            bool any = false, c[W];
            for (int j = 0; j < W; ++j) {
                val[j] = val[j] * u[j];
                c[j] = (nem[j] < val[j]) && valid[j];
                any |= c[j];
            }
            if (!any) break;
            for (int j = 0; j < W; ++j) res[j] += c[j] ? 1 : 0;
        }
        for (int j = 0; j < W; ++j) valid_res[j] = (float)res[j];
    }
    if (!all_valid) {
        for (;;) {
            float u1[W], u2[W], fres[W], p12[W];
            bool c1[W];
            urW<W>(e, u1);
            for (int j = 0; j < W; ++j) {
                float s, c;
                psincos(kPiF * u1[j], s, c);
                float ys = s / c;
                fres[j] = ptruncate(psq[j] * ys + pmean[j]);
                float pp1 = (F(1.2) * psq[j]) * (1.0f + ys * ys);
                float pp2 = pexp((((pg1[j] - plgamma_approx(fres[j] + 1.0f)) -
                                   plgamma_approx((ptr[j] + 1.0f) - fres[j])) +
                                  fres[j] * plp[j]) +
                                 (ptr[j] - fres[j]) * plq[j]);
                c1[j] = (0.0f <= fres[j]) && (fres[j] <= ptr[j]);
                p12[j] = pp1 * pp2;
            }
            urW<W>(e, u2);
            bool all = true;
            for (int j = 0; j < W; ++j) {
                bool fv = c1[j] && (u2[j] <= p12[j]);
                if (fv && !valid[j]) valid_res[j] = fres[j];
                valid[j] = valid[j] || fv;
                all = all && valid[j];
            }
            if (all) break;
        }
    }
    for (int j = 0; j < W; ++j) {
        int32_t k = (int32_t)valid_res[j];
        out[j] = (p[j] == psmall[j]) ? k : trials[j] - k;
    }
}

// --------------------------------------------------------------- helpers ---
template <class Fn>
void for_engines(int64_t n_engines, int threads, Fn fn) {
    if (threads <= 1 || n_engines <= 1) {
        for (int64_t g = 0; g < n_engines; ++g) fn(g);
        return;
    }
    std::vector<std::thread> ts;
    for (int t = 0; t < threads; ++t) {
        ts.emplace_back([=, &fn]() {
            for (int64_t g = t; g < n_engines; g += threads) fn(g);
        });
    }
    for (auto& th : ts) th.join();
}

}  // namespace

extern "C" {

int orc_state_bytes() { return (int)sizeof(Vmt); }

void orc_seed(void* states, const uint64_t* seeds, int64_t n_engines) {
    Vmt* e = (Vmt*)states;
    for (int64_t g = 0; g < n_engines; ++g) vmt_seed(e[g], seeds[g]);
}

// Raw u64 draws: out[g*calls*4 + 4k + l] = lane l of call k of engine g.
void orc_fill_u64(void* states, int64_t n_engines, uint64_t* out, int64_t calls, int threads) {
    Vmt* e = (Vmt*)states;
    for_engines(n_engines, threads, [&](int64_t g) {
        uint64_t* o = out + g * calls * 4;
        for (int64_t k = 0; k < calls; ++k) vmt_draw(e[g], o + 4 * k);
    });
}

// MersenneTwister::discard (PacketRandomEngine.h:317-323): `calls` draws.
void orc_discard(void* states, int64_t n_engines, int64_t calls, int threads) {
    Vmt* e = (Vmt*)states;
    for_engines(n_engines, threads, [&](int64_t g) {
        uint64_t w[4];
        for (int64_t k = 0; k < calls; ++k) vmt_draw(e[g], w);
    });
}

// P8_mt19937_64 / ParallelRandomEngineAdaptor<uint64_t, Vmt19937_64, 8>
// (PacketRandomEngine.h:429-517, 600-609): two engines seeded s and s+4,
// buffer = [engine0 call, engine1 call]; out[8k + 4i + l].
void orc_fill_p8_u64(void* two_states, uint64_t* out, int64_t n_refills) {
    Vmt* e = (Vmt*)two_states;
    for (int64_t k = 0; k < n_refills; ++k) {
        vmt_draw(e[0], out + 8 * k);
        vmt_draw(e[1], out + 8 * k + 4);
    }
}

// half() draws: out[g*halves*2 + 2h + {0,1}].
void orc_fill_half_u64(void* states, int64_t n_engines, uint64_t* out, int64_t halves) {
    Vmt* e = (Vmt*)states;
    for (int64_t g = 0; g < n_engines; ++g)
        for (int64_t h = 0; h < halves; ++h) vmt_half(e[g], out + (g * halves + h) * 2);
}

// uniformReal: StdUniformRealGen (Basic.h:584-590) when (a,b)=(0,1);
// UniformRealGen (Basic.h:626-634) otherwise: u*(b-a)+a.
void orc_fill_uniform_f32(void* states, int64_t n_engines, float* out, int64_t per_engine,
                          float a, float b, int affine, int threads) {
    Vmt* e = (Vmt*)states;
    const float slope = b - a;
    for_engines(n_engines, threads, [&](int64_t g) {
        float* o = out + g * per_engine;
        float u[8];
        for (int64_t i = 0; i < per_engine; i += 8) {
            ur8(e[g], u);
            for (int j = 0; j < 8; ++j) o[i + j] = affine ? u[j] * slope + a : u[j];
        }
    });
}

// uniformReal<double>: UniformRealUtils<Packet4d>::uniform_real
// (arch/AVX/RandUtils.h:212-225) = bit_to_ur_double of one full call.
void orc_fill_uniform_f64(void* states, int64_t n_engines, double* out, int64_t per_engine,
                          double a, double b, int affine, int threads) {
    Vmt* e = (Vmt*)states;
    const double slope = b - a;
    for_engines(n_engines, threads, [&](int64_t g) {
        double* o = out + g * per_engine;
        uint64_t p[4];
        for (int64_t i = 0; i < per_engine; i += 4) {
            vmt_draw(e[g], p);
            for (int j = 0; j < 4; ++j) {
                double u = bit_to_ur_double(p[j]);
                o[i + j] = affine ? u * slope + a : u;
            }
        }
    });
}

// normal<double>: StdNormalGen at Packet4d (NormalExp.h:60-82), NormalGen
// z*sd+mean when affine=1 (127-135). Draws 2m, 2m+1 -> four pairs: the sine
// packet is returned, the cosine packet cached for the next packetOp; a fresh
// generator per engine block drops a trailing cosine packet.
void orc_fill_normal_f64(void* states, int64_t n_engines, double* out, int64_t per_engine,
                         double mean, double sd, int affine, int threads) {
    Vmt* e = (Vmt*)states;
    for_engines(n_engines, threads, [&](int64_t g) {
        double* o = out + g * per_engine;
        double cache[4];
        bool valid = false;
        uint64_t p1[4], p2[4];
        for (int64_t i = 0; i < per_engine; i += 4) {
            if (valid) {
                for (int j = 0; j < 4; ++j) o[i + j] = cache[j];
                valid = false;
            } else {
                vmt_draw(e[g], p1);
                vmt_draw(e[g], p2);
                for (int j = 0; j < 4; ++j) {
                    const double u1 = 1.0 - bit_to_ur_double(p1[j]);
                    const double u2 = bit_to_ur_double(p2[j]);
                    const double radius = std::sqrt(-2.0 * plog_d(u1));
                    const double theta = (2 * kPi) * u2;
                    double sn, cs;
                    psincos_d(theta, sn, cs);
                    cache[j] = radius * cs;
                    o[i + j] = radius * sn;
                }
                valid = true;
            }
            if (affine) for (int j = 0; j < 4; ++j) o[i + j] = o[i + j] * sd + mean;
        }
    });
}

// ExponentialGen<double> (NormalExp.h:391-398): -(log(1-u)/lambda).
void orc_fill_exponential_f64(void* states, int64_t n_engines, double* out, int64_t per_engine,
                              double lambda, int threads) {
    Vmt* e = (Vmt*)states;
    for_engines(n_engines, threads, [&](int64_t g) {
        double* o = out + g * per_engine;
        uint64_t p[4];
        for (int64_t i = 0; i < per_engine; i += 4) {
            vmt_draw(e[g], p);
            for (int j = 0; j < 4; ++j) o[i + j] = -(plog_d(1.0 - bit_to_ur_double(p[j])) / lambda);
        }
    });
}

// CauchyGen<double> (NormalExp.h:843-855): a + b*(sin/cos)(pi*(u-0.5)).
void orc_fill_cauchy_f64(void* states, int64_t n_engines, double* out, int64_t per_engine,
                         double a, double b, int threads) {
    Vmt* e = (Vmt*)states;
    for_engines(n_engines, threads, [&](int64_t g) {
        double* o = out + g * per_engine;
        uint64_t p[4];
        for (int64_t i = 0; i < per_engine; i += 4) {
            vmt_draw(e[g], p);
            for (int j = 0; j < 4; ++j) {
                double sn, cs;
                psincos_d(kPi * (bit_to_ur_double(p[j]) - 0.5), sn, cs);
                o[i + j] = a + b * (sn / cs);
            }
        }
    });
}

// normal: StdNormalGen when affine=0, NormalGen z*sd+mean (NormalExp.h:
// 127-135) when affine=1. A fresh generator per engine block.
void orc_fill_normal_f32(void* states, int64_t n_engines, float* out, int64_t per_engine,
                         float mean, float sd, int affine, int threads) {
    Vmt* e = (Vmt*)states;
    for_engines(n_engines, threads, [&](int64_t g) {
        float* o = out + g * per_engine;
        StdNormal8 gen;
        for (int64_t i = 0; i < per_engine; i += 8) {
            gen.op(e[g], o + i);
            if (affine) for (int j = 0; j < 8; ++j) o[i + j] = o[i + j] * sd + mean;
        }
    });
}

// NormalVGen (NormalExp.h:138-163): z*sd[i]+mean[i].
void orc_fill_normal_vop_f32(void* states, int64_t n_engines, float* out, int64_t per_engine,
                             const float* mean, const float* sd, int threads) {
    Vmt* e = (Vmt*)states;
    for_engines(n_engines, threads, [&](int64_t g) {
        float* o = out + g * per_engine;
        const float* m = mean + g * per_engine;
        const float* s = sd + g * per_engine;
        StdNormal8 gen;
        for (int64_t i = 0; i < per_engine; i += 8) {
            gen.op(e[g], o + i);
            for (int j = 0; j < 8; ++j) o[i + j] = o[i + j] * s[i + j] + m[i + j];
        }
    });
}

// ExponentialGen (NormalExp.h:391-398): -(log(1-u)/lambda); pnegate flips the sign bit.
void orc_fill_exponential_f32(void* states, int64_t n_engines, float* out, int64_t per_engine,
                              float lambda, int threads) {
    Vmt* e = (Vmt*)states;
    for_engines(n_engines, threads, [&](int64_t g) {
        float* o = out + g * per_engine;
        float u[8];
        for (int64_t i = 0; i < per_engine; i += 8) {
            ur8(e[g], u);
            for (int j = 0; j < 8; ++j) o[i + j] = -(plog(1.0f - u[j]) / lambda);
        }
    });
}

// CauchyGen (NormalExp.h:843-855): a + b*tan(pi*(u-0.5)) via psincos.
// LognormalGen (NormalExp.h:171-210): pexp of NormalGen's packets.
void orc_fill_lognormal_f32(void* states, int64_t n_engines, float* out, int64_t per_engine,
                            float mean, float sd, int threads) {
    Vmt* e = (Vmt*)states;
    for_engines(n_engines, threads, [&](int64_t g) {
        float* o = out + g * per_engine;
        StdNormal8 gen;
        for (int64_t i = 0; i < per_engine; i += 8) {
            gen.op(e[g], o + i);
            for (int j = 0; j < 8; ++j) o[i + j] = pexp(o[i + j] * sd + mean);
        }
    });
}

// WeibullGen (NormalExp.h:649-655): b * exp(log(-log(1 - U)) * (1/a)).
void orc_fill_weibull_f32(void* states, int64_t n_engines, float* out, int64_t per_engine,
                          float a, float b, int threads) {
    Vmt* e = (Vmt*)states;
    const float inv_a = 1 / a;
    for_engines(n_engines, threads, [&](int64_t g) {
        float* o = out + g * per_engine;
        float u[8];
        for (int64_t i = 0; i < per_engine; i += 8) {
            ur8(e[g], u);
            for (int j = 0; j < 8; ++j) o[i + j] = pexp(plog(-plog(1.0f - u[j])) * inv_a) * b;
        }
    });
}

// ExtremeValueGen (NormalExp.h:725-731): a - log(-log(U + eps/8)) * b.
void orc_fill_extreme_value_f32(void* states, int64_t n_engines, float* out, int64_t per_engine,
                                float a, float b, int threads) {
    Vmt* e = (Vmt*)states;
    for_engines(n_engines, threads, [&](int64_t g) {
        float* o = out + g * per_engine;
        float u[8];
        for (int64_t i = 0; i < per_engine; i += 8) {
            nzur8(e[g], u);
            for (int j = 0; j < 8; ++j) o[i + j] = a - plog(-plog(u[j])) * b;
        }
    });
}

// BernoulliGen<float>::packetOp (Basic.h:693-707) at Packet8f: 31 bits of
// each 32-bit word against p = uint32(p * 2^31), compared as signed int32
// lanes (pcmplt on Packet8i). vp != NULL: BernoulliVGen<float> (Basic.h:
// 726-733), p per element = cvttps(p_i * 2^31).
void orc_fill_bernoulli_f32(void* states, int64_t n_engines, float* out, int64_t per_engine, double p,
                            const float* vp, int threads) {
    Vmt* e = (Vmt*)states;
    const int32_t pi = (int32_t)(uint32_t)(p * 0x80000000);
    for_engines(n_engines, threads, [&](int64_t g) {
        float* o = out + g * per_engine;
        uint64_t w[4];
        for (int64_t i = 0; i < per_engine; i += 8) {
            vmt_draw(e[g], w);
            for (int j = 0; j < 8; ++j) {
                const int32_t r = (int32_t)((uint32_t)(w[j >> 1] >> (32 * (j & 1))) & 0x7FFFFFFFu);
                const int32_t t = vp ? cvttps(vp[g * per_engine + i + j] * (float)0x80000000) : pi;
                o[i + j] = r < t ? 1.0f : 0.0f;
            }
        }
    });
}

// Balanced2Gen (Basic.h:437-470) and RandUtils::balanced (RandUtils.h:42-45):
// zero_to_one * slope + bias with slope = b - a, bias = a (float); va != NULL:
// BalancedVGen (Basic.h:487-493), zero_to_one * (b_i - a_i) + a_i.
void orc_fill_balanced_f32(void* states, int64_t n_engines, float* out, int64_t per_engine, float a, float b,
                           const float* va, const float* vb, int threads) {
    Vmt* e = (Vmt*)states;
    const float slope = b - a, bias = a;
    for_engines(n_engines, threads, [&](int64_t g) {
        float* o = out + g * per_engine;
        uint64_t w[4];
        for (int64_t i = 0; i < per_engine; i += 8) {
            vmt_draw(e[g], w);
            for (int j = 0; j < 8; ++j) {
                const float z = zero_to_one_float((uint32_t)(w[j >> 1] >> (32 * (j & 1))));
                if (va) {
                    const int64_t k = g * per_engine + i + j;
                    o[i + j] = z * (vb[k] - va[k]) + va[k];
                } else {
                    o[i + j] = z * slope + bias;
                }
            }
        }
    });
}

// Vectorize-over-params single-draw generators (kind numbers of the device
// library's VopKind): 0 UniformRealVGen (Basic.h:653-660), 1 ExponentialVGen
// (NormalExp.h:418-424), 2 CauchyVGen (NormalExp.h:875-884), 3 WeibullVGen
// (NormalExp.h:675-681), 4 ExtremeValueVGen (NormalExp.h:752-760), 5
// StudentTVGen (NormalExp.h:332-345, two draws per packet), 6 LognormalVGen
// (NormalExp.h:229-233) = pexp of NormalVGen.
void orc_fill_vop_f32(void* states, int64_t n_engines, int kind, float* out, int64_t per_engine,
                      const float* va, const float* vb, int threads) {
    Vmt* e = (Vmt*)states;
    if (kind == 6) {
        orc_fill_normal_vop_f32(states, n_engines, out, per_engine, va, vb, threads);
        for (int64_t i = 0; i < n_engines * per_engine; ++i) out[i] = pexp(out[i]);
        return;
    }
    for_engines(n_engines, threads, [&](int64_t g) {
        float* o = out + g * per_engine;
        float u[8], u2[8];
        for (int64_t i = 0; i < per_engine; i += 8) {
            if (kind == 4) nzur8(e[g], u);
            else ur8(e[g], u);
            if (kind == 5) ur8(e[g], u2);
            for (int j = 0; j < 8; ++j) {
                const int64_t k = g * per_engine + i + j;
                const float a = va[k];
                float v;
                if (kind == 0) {
                    v = u[j] * (vb[k] - a) + a;
                } else if (kind == 1) {
                    v = -(plog(1.0f - u[j]) / a);
                } else if (kind == 2) {
                    float sn, cs;
                    psincos(kPiF * (u[j] - 0.5f), sn, cs);
                    v = a + vb[k] * (sn / cs);
                } else if (kind == 3) {
                    v = pexp(plog(-plog(1.0f - u[j])) * (1.0f / a)) * vb[k];
                } else if (kind == 4) {
                    v = a - plog(-plog(u[j])) * vb[k];
                } else {
                    const float u1 = 1.0f - u[j];
                    const float radius = std::sqrt(a * (pexp(plog(u1) * (-2.0f / a)) - 1.0f));
                    float sn, cs;
                    psincos((float)(2 * kPi) * u2[j], sn, cs);
                    v = radius * sn;
                }
                o[i + j] = v;
            }
        }
    });
}

// Packet4i fills through half(): BernoulliGen<int32_t> (Basic.h:693-707)
// mode 0, GeometricGen (Discrete.h:1372-1380) mode 1 with rlog_q = 1/log(1-p)
// in double then float, RandbitsGen<int32_t> (Basic.h:394-400) mode 2.
void orc_fill_half_i32(void* states, int64_t n_engines, int mode, int32_t* out, int64_t per_engine, double p,
                       int threads) {
    Vmt* e = (Vmt*)states;
    const int32_t pi = (int32_t)(uint32_t)(p * 0x80000000);
    const float rlog_q = (float)(1 / std::log(1 - p));
    for_engines(n_engines, threads, [&](int64_t g) {
        int32_t* o = out + g * per_engine;
        uint64_t h[2];
        for (int64_t i = 0; i < per_engine; i += 4) {
            vmt_half(e[g], h);
            for (int j = 0; j < 4; ++j) {
                const uint32_t w = (uint32_t)(h[j >> 1] >> (32 * (j & 1)));
                if (mode == 0) o[i + j] = (int32_t)(w & 0x7FFFFFFFu) < pi ? 1 : 0;
                else if (mode == 1) o[i + j] = cvttps(ptruncate(plog(1.0f - bit_to_ur_float(w)) * rlog_q));
                else o[i + j] = (int32_t)w;
            }
        }
    });
}

// UniformIntGen<int32_t> (Discrete.h:215-335) at Packet4i; the constructor
// (229-246) for ranges with 0 <= max - min < 2^31: bitsize = ceil(log2(pdiff
// + 1)), bitmask = 2^bitsize - 1. pdiff = 0: the constant min, no draws;
// pdiff = bitmask: (rx & bitmask) + min per half draw; otherwise candidates
// rx & bitmask, rx >>= bitsize while bitcnt + bitsize < 32, through
// CompressMask<16> with a signed compare against pdiff + 1; the rest of a
// half draw is dropped once a packet is emitted. A fresh generator per fill.
void orc_fill_uniform_int_i32(void* states, int64_t n_engines, int32_t* out, int64_t per_engine, int32_t lo,
                              int32_t hi, int threads) {
    Vmt* e = (Vmt*)states;
    const uint64_t pdiff = (uint64_t)((int64_t)hi - (int64_t)lo);
    const uint64_t bitsize = pdiff == 0 ? 0 : (uint64_t)std::ceil(std::log2((double)(pdiff + 1)));
    const uint64_t bitmask = bitsize == 0 ? 0 : (~0ull >> (64 - bitsize));
    for_engines(n_engines, threads, [&](int64_t g) {
        int32_t* o = out + g * per_engine;
        int32_t rest[4];
        int rest_cnt = 0;
        for (int64_t i = 0; i < per_engine; i += 4) {
            if (pdiff == 0) {
                for (int j = 0; j < 4; ++j) o[i + j] = lo;
                continue;
            }
            uint64_t h[2];
            vmt_half(e[g], h);
            uint32_t rx[4];
            for (int j = 0; j < 4; ++j) rx[j] = (uint32_t)(h[j >> 1] >> (32 * (j & 1)));
            if (pdiff == bitmask) {
                for (int j = 0; j < 4; ++j) o[i + j] = (int32_t)((rx[j] & (uint32_t)bitmask) + (uint32_t)lo);
                continue;
            }
            const int32_t plen = (int32_t)(uint32_t)(pdiff + 1);
            uint64_t bitcnt = bitsize;
            for (;;) {
                int32_t cands[4];
                bool m[4];
                for (int j = 0; j < 4; ++j) {
                    cands[j] = (int32_t)(rx[j] & (uint32_t)bitmask);
                    m[j] = cands[j] < plen;
                }
                if (compress_append<4, int32_t>(cands, m, rest, rest_cnt)) {
                    for (int j = 0; j < 4; ++j) o[i + j] = (int32_t)((uint32_t)cands[j] + (uint32_t)lo);
                    break;
                }
                if (bitcnt + bitsize < 32) {
                    for (int j = 0; j < 4; ++j) rx[j] >>= bitsize;
                    bitcnt += bitsize;
                } else {
                    vmt_half(e[g], h);
                    for (int j = 0; j < 4; ++j) rx[j] = (uint32_t)(h[j >> 1] >> (32 * (j & 1)));
                    bitcnt = bitsize;
                }
            }
        }
    });
}

// DiscreteGen<int32_t, float> (Discrete.h:486-603) at Packet4i: K < 16: a
// float CDF (cdf[i] = float(c_i / c_K)) counted against a half draw's
// uniform; otherwise AliasMethod<float, int32_t> (Discrete.h:23-207 with the
// float branch of buildTable), index bits from one half draw, the uniform
// from the next. Tables come from orc_alias_table_f32.
void orc_fill_discrete_f_i32(void* states, int64_t n_engines, int32_t* out, int64_t per_engine, const float* cdf,
                             int32_t k, const float* prob, const int32_t* alias, uint32_t bitmask) {
    Vmt* e = (Vmt*)states;
    for (int64_t g = 0; g < n_engines; ++g) {
        int32_t* o = out + g * per_engine;
        for (int64_t i = 0; i < per_engine; i += 4) {
            uint64_t h[2];
            vmt_half(e[g], h);
            if (cdf) {
                for (int j = 0; j < 4; ++j) {
                    const float u = bit_to_ur_float((uint32_t)(h[j >> 1] >> (32 * (j & 1))));
                    int32_t r = k;
                    for (int c = 0; c < k; ++c) r -= u < cdf[c] ? 1 : 0;
                    o[i + j] = r;
                }
            } else {
                uint32_t idx[4];
                for (int j = 0; j < 4; ++j) idx[j] = (uint32_t)(h[j >> 1] >> (32 * (j & 1))) & bitmask;
                vmt_half(e[g], h);
                for (int j = 0; j < 4; ++j) {
                    const float u = bit_to_ur_float((uint32_t)(h[j >> 1] >> (32 * (j & 1))));
                    o[i + j] = u < prob[idx[j]] ? (int32_t)idx[j] : alias[idx[j]];
                }
            }
        }
    }
}

void orc_fill_cauchy_f32(void* states, int64_t n_engines, float* out, int64_t per_engine,
                         float a, float b, int threads) {
    Vmt* e = (Vmt*)states;
    for_engines(n_engines, threads, [&](int64_t g) {
        float* o = out + g * per_engine;
        float u[8];
        for (int64_t i = 0; i < per_engine; i += 8) {
            ur8(e[g], u);
            for (int j = 0; j < 8; ++j) {
                float s, c;
                psincos(kPiF * (u[j] - F(0.5)), s, c);
                o[i + j] = a + b * (s / c);
            }
        }
    });
}

// StudentTGen (NormalExp.h:285-301). The reference calls Eigen's psin; the
// sine of the reference's own psincos stands in for it here and on the GPU.
void orc_fill_studentt_f32(void* states, int64_t n_engines, float* out, int64_t per_engine,
                           float n, int threads) {
    Vmt* e = (Vmt*)states;
    const float k = -2 / n;
    for_engines(n_engines, threads, [&](int64_t g) {
        float* o = out + g * per_engine;
        float u1[8], u2[8];
        for (int64_t i = 0; i < per_engine; i += 8) {
            ur8(e[g], u1);
            ur8(e[g], u2);
            for (int j = 0; j < 8; ++j) {
                float a = 1.0f - u1[j];
                float radius = std::sqrt(n * (pexp(plog(a) * k) - 1.0f));
                float theta = F(2 * kPi) * u2[j];
                float s, c;
                psincos(theta, s, c);
                o[i + j] = radius * s;
            }
        }
    });
}

// GammaGen; ChiSquaredGen(n) = GammaGen(n*0.5, 2) (NormalExp.h:781-784).
void orc_fill_gamma_f32(void* states, int64_t n_engines, float* out, int64_t per_engine,
                        float alpha, float beta, int threads) {
    Vmt* e = (Vmt*)states;
    for_engines(n_engines, threads, [&](int64_t g) {
        float* o = out + g * per_engine;
        Gamma8 gen(alpha, beta);
        for (int64_t i = 0; i < per_engine; i += 8) gen.op(e[g], o + i);
    });
}

// BetaGen (NormalExp.h:896-985) at Packet8f; ratio != 0: FisherFGen
// (NormalExp.h:987-1023) with ratio = b/a, x -> ratio * x / (1 - x).
// a < 1 and b < 1: Johnk through CompressMask<32>; otherwise GammaGen(a) and
// GammaGen(b) (beta 1), one packet each in turn, p1 / (p1 + p2).
void orc_fill_beta_f32(void* states, int64_t n_engines, float* out, int64_t per_engine, float a, float b,
                       float ratio, int threads) {
    Vmt* e = (Vmt*)states;
    auto fin = [&](float x) { return ratio != 0.0f ? (ratio * x) / (1.0f - x) : x; };
    for_engines(n_engines, threads, [&](int64_t g) {
        float* o = out + g * per_engine;
        if (a < 1 && b < 1) {
            float rest[8];
            int rest_cnt = 0;
            for (int64_t i = 0; i < per_engine; i += 8) {
                for (;;) {
                    float u1[8], u2[8], cands[8];
                    bool acc[8];
                    ur8(e[g], u1);
                    ur8(e[g], u2);
                    for (int j = 0; j < 8; ++j) {
                        const float p1 = pexp(plog(u1[j]) * (1 / a));
                        const float p2 = pexp(plog(u2[j]) * (1 / b));
                        const float x = p1 + p2;
                        cands[j] = p1 / x;
                        acc[j] = x <= 1.0f;
                    }
                    if (compress_append<8>(cands, acc, rest, rest_cnt)) {
                        for (int j = 0; j < 8; ++j) o[i + j] = fin(cands[j]);
                        break;
                    }
                }
            }
            return;
        }
        Gamma8 g1(a, 1.0f), g2(b, 1.0f);
        for (int64_t i = 0; i < per_engine; i += 8) {
            float p1[8], p2[8];
            g1.op(e[g], p1);
            g2.op(e[g], p2);
            for (int j = 0; j < 8; ++j) o[i + j] = fin(p1[j] / (p1[j] + p2[j]));
        }
    });
}

// NegativeBinomialGen<int32_t>::packetOp (GammaPoisson.h:29-100) at
// Packet4i: mean = GammaGen<float>(trials, (1 - p) / p) at Packet4f (half
// draws, queue Q); if no lane's mean exceeds 12, Knuth's product loop with
// per-lane thresholds e^-mean; otherwise tangent rejection with per-lane
// constants whose candidates go through the gamma's own queue Q, emitted
// values cast with cvttps. A fresh generator per engine fill.
void orc_fill_negative_binomial_i32(void* states, int64_t n_engines, int32_t* out, int64_t per_engine,
                                    int32_t trials, double p, int threads) {
    Vmt* e = (Vmt*)states;
    for_engines(n_engines, threads, [&](int64_t g) {
        int32_t* o = out + g * per_engine;
        GammaPk<4> gam((float)trials, (float)((1 - p) / p));
        for (int64_t i = 0; i < per_engine; i += 4) {
            float mean[4];
            gam.op(e[g], mean);
            bool any_big = false;
            for (int j = 0; j < 4; ++j) any_big |= 12.0f < mean[j];
            if (!any_big) {
                float val[4] = {1.0f, 1.0f, 1.0f, 1.0f}, ne[4];
                int32_t res[4] = {0, 0, 0, 0};
                for (int j = 0; j < 4; ++j) ne[j] = pexp(-mean[j]);
                for (;;) {
                    float u[4];
                    ur4(e[g], u);
                    bool anyc = false;
                    for (int j = 0; j < 4; ++j) {
                        val[j] = val[j] * u[j];
                        const bool c = ne[j] < val[j];
                        anyc |= c;
                        res[j] += c ? 1 : 0;
                    }
                    if (!anyc) break;
                }
                for (int j = 0; j < 4; ++j) o[i + j] = res[j];
                continue;
            }
            float sq[4], lm[4], g1[4];
            for (int j = 0; j < 4; ++j) {
                sq[j] = std::sqrt(2.0f * mean[j]);
                lm[j] = plog(mean[j]);
                g1[j] = mean[j] * lm[j] - plgamma_approx(mean[j] + 1.0f);
            }
            for (;;) {
                float u[4], u2[4], cands[4];
                bool acc[4];
                ur4(e[g], u);
                float yx[4];
                for (int j = 0; j < 4; ++j) {
                    float sn, cs;
                    psincos(kPiF * u[j], sn, cs);
                    yx[j] = sn / cs;
                    cands[j] = ptruncate(sq[j] * yx[j] + mean[j]);
                }
                ur4(e[g], u2);
                for (int j = 0; j < 4; ++j) {
                    const float p1 = (yx[j] * yx[j] + 1.0f) * 0.9f;
                    const float p2 = pexp((cands[j] * lm[j] - plgamma_approx(cands[j] + 1.0f)) - g1[j]);
                    acc[j] = (0.0f <= cands[j]) && (u2[j] <= p1 * p2);
                }
                if (compress_append<4>(cands, acc, gam.rest, gam.rest_cnt)) {
                    for (int j = 0; j < 4; ++j) o[i + j] = cvttps(cands[j]);
                    break;
                }
            }
        }
    });
}

void orc_fill_poisson_i32(void* states, int64_t n_engines, int32_t* out, int64_t per_engine,
                          double mean, int threads) {
    Vmt* e = (Vmt*)states;
    for_engines(n_engines, threads, [&](int64_t g) {
        int32_t* o = out + g * per_engine;
        Poisson4 gen(mean);
        for (int64_t i = 0; i < per_engine; i += 4) gen.op(e[g], o + i);
    });
}

void orc_fill_binomial_i32(void* states, int64_t n_engines, int32_t* out, int64_t per_engine,
                           int32_t trials, double p, int threads) {
    Vmt* e = (Vmt*)states;
    for_engines(n_engines, threads, [&](int64_t g) {
        int32_t* o = out + g * per_engine;
        Binomial4 gen(trials, p);
        for (int64_t i = 0; i < per_engine; i += 4) gen.op(e[g], o + i);
    });
}

void orc_fill_binomial_vop_i32(void* states, int64_t n_engines, int32_t* out, int64_t per_engine,
                               const int32_t* trials, const float* p, int threads) {
    Vmt* e = (Vmt*)states;
    for_engines(n_engines, threads, [&](int64_t g) {
        int64_t base = g * per_engine;
        for (int64_t i = 0; i < per_engine; i += 4)
            binomialv<4>(e[g], trials + base + i, p + base + i, out + base + i);
    });
}

// binomial(urng, trials, p) with BinomialVGen<float> (Discrete.h:1196-1331):
// Packet8f lanes on full draws, per-element trials and p.
void orc_fill_binomial_vopf_i32(void* states, int64_t n_engines, int32_t* out, int64_t per_engine,
                                const int32_t* trials, const float* p, int threads) {
    Vmt* e = (Vmt*)states;
    for_engines(n_engines, threads, [&](int64_t g) {
        int64_t base = g * per_engine;
        for (int64_t i = 0; i < per_engine; i += 8)
            binomialv<8>(e[g], trials + base + i, p + base + i, out + base + i);
    });
}

// Reused generator objects (`auto gen = Rand::makeXxxGen(...)` followed by
// several gen.generate() calls): per-engine generators that keep the normal
// generator's cached cos packet and the rejection samplers' CompressMask
// leftovers between fills (EigenRand/Dists/Basic.h:268-342). dist numbers are
// erd_dist's: 3 std normal, 4 normal(a, b), 9 gamma(a, b), 10 Poisson(a),
// 11 binomial(a, b), 16 lognormal(a, b); other kinds have no state and use
// the plain fills.
struct GenObjects {
    int dist;
    double a, b;
    std::vector<StdNormal8> normal;
    std::vector<Gamma8> gamma;
    std::vector<Poisson4> poisson;
    std::vector<Binomial4> binomial;
};

void* orc_gen_create(int dist, double a, double b, int64_t n) {
    GenObjects* h = new GenObjects{dist, a, b, {}, {}, {}, {}};
    for (int64_t g = 0; g < n; ++g) {
        if (dist == 3 || dist == 4 || dist == 16) h->normal.emplace_back();
        if (dist == 9) h->gamma.emplace_back((float)a, (float)b);
        if (dist == 10) h->poisson.emplace_back(a);
        if (dist == 11) h->binomial.emplace_back((int32_t)a, b);
    }
    return h;
}

void orc_gen_destroy(void* h) { delete (GenObjects*)h; }

// gen.generate(per_engine, 1, engine_g) for every engine g.
void orc_gen_fill(void* handle, void* states, int64_t n_engines, void* out, int64_t per_engine, int threads) {
    GenObjects* h = (GenObjects*)handle;
    Vmt* e = (Vmt*)states;
    for_engines(n_engines, threads, [&](int64_t g) {
        if (h->dist == 3 || h->dist == 4 || h->dist == 16) {
            float* o = (float*)out + g * per_engine;
            for (int64_t i = 0; i < per_engine; i += 8) {
                h->normal[g].op(e[g], o + i);
                if (h->dist == 4)
                    for (int j = 0; j < 8; ++j) o[i + j] = o[i + j] * (float)h->b + (float)h->a;
                if (h->dist == 16)
                    for (int j = 0; j < 8; ++j) o[i + j] = pexp(o[i + j] * (float)h->b + (float)h->a);
            }
        } else if (h->dist == 9) {
            float* o = (float*)out + g * per_engine;
            for (int64_t i = 0; i < per_engine; i += 8) h->gamma[g].op(e[g], o + i);
        } else if (h->dist == 10) {
            int32_t* o = (int32_t*)out + g * per_engine;
            for (int64_t i = 0; i < per_engine; i += 4) h->poisson[g].op(e[g], o + i);
        } else if (h->dist == 11) {
            int32_t* o = (int32_t*)out + g * per_engine;
            for (int64_t i = 0; i < per_engine; i += 4) h->binomial[g].op(e[g], o + i);
        }
    });
}

// Elementwise access to the packet math, for pinning it against libm.
void orc_plog(const float* x, float* y, int64_t n) { for (int64_t i = 0; i < n; ++i) y[i] = plog(x[i]); }
void orc_pexp(const float* x, float* y, int64_t n) { for (int64_t i = 0; i < n; ++i) y[i] = pexp(x[i]); }
void orc_psincos(const float* x, float* s, float* c, int64_t n) {
    for (int64_t i = 0; i < n; ++i) psincos(x[i], s[i], c[i]);
}
void orc_plgamma(const float* x, float* y, int64_t n) { for (int64_t i = 0; i < n; ++i) y[i] = plgamma_approx(x[i]); }
void orc_bit_to_ur_float(const uint32_t* w, float* y, int64_t n) { for (int64_t i = 0; i < n; ++i) y[i] = bit_to_ur_float(w[i]); }
void orc_plog_f64(const double* x, double* y, int64_t n) { for (int64_t i = 0; i < n; ++i) y[i] = plog_d(x[i]); }
void orc_psincos_f64(const double* x, double* s, double* c, int64_t n) {
    for (int64_t i = 0; i < n; ++i) psincos_d(x[i], s[i], c[i]);
}

}  // extern "C"

extern "C" {
// Runs the CompressMask queue over candidate packets (test hook): values and
// masks are n_cand x W; writes up to max_packets emitted packets to out and
// returns the number of candidates consumed when the last one was emitted.
int64_t orc_compress_stream(int W, const float* values, const uint8_t* masks, int64_t n_cand,
                            float* out, int64_t max_packets) {
    float rest[8];
    int rest_cnt = 0;
    int64_t emitted = 0;
    for (int64_t c = 0; c < n_cand && emitted < max_packets; ++c) {
        float v[8];
        bool m[8];
        for (int j = 0; j < W; ++j) { v[j] = values[c * W + j]; m[j] = masks[c * W + j] != 0; }
        bool full = (W == 8) ? compress_append<8>(v, m, rest, rest_cnt) : compress_append<4>(v, m, rest, rest_cnt);
        if (full) {
            for (int j = 0; j < W; ++j) out[emitted * W + j] = v[j];
            if (++emitted == max_packets) return c + 1;
        }
    }
    return -1;
}
}

// ----------------------------------------------------------------------------
// Vectorize-over-params gamma and Poisson. The reference has no GammaVGen or
// PoissonVGen; these follow the design of its BinomialVGen<int32_t>
// (Discrete.h:1124-1192): per-lane parameters computed in float, lockstep
// iterations, each lane keeps its first accepted candidate.
namespace {

// Gamma: Packet8f (full draws). Every iteration draws three packets (ux, vx,
// u3 as in the GS branch, NormalExp.h:536-558); lanes with alpha <= 1 use
// them as GS does, lanes with alpha > 1 run the tangent step (580-603) on the
// first two. Output beta * x.
void gammav8(Vmt& e, const float alpha[8], const float beta[8], float out[8]) {
    float px[8], inv[8], am1[8], sq[8], res[8] = {0};
    bool done[8];
    for (int j = 0; j < 8; ++j) {
        px[j] = F(kE) / (alpha[j] + F(kE));
        inv[j] = 1.0f / alpha[j];
        am1[j] = alpha[j] - 1.0f;
        sq[j] = std::sqrt(2.0f * alpha[j] - 1.0f);
        done[j] = false;
    }
    for (;;) {
        float u1[8], u2[8], u3[8];
        ur8(e, u1);
        ur8(e, u2);
        ur8(e, u3);
        bool all = true;
        for (int j = 0; j < 8; ++j) {
            float x;
            bool acc;
            if (alpha[j] <= 1.0f) {
                float vx = u2[j] + 1.1920928955078125e-07f / 8;
                float lv = plog(vx);
                float x1 = pexp(inv[j] * lv);
                float q1 = pexp(-x1);
                float x2 = 1.0f - lv;
                float q2 = pexp(plog(x2) * am1[j]);
                bool c = u1[j] < px[j];
                x = c ? x1 : x2;
                acc = u3[j] < (c ? q1 : q2);
            } else {
                float ys, yc;
                psincos(kPiF * u1[j], ys, yc);
                float yx = ys / yc;
                x = sq[j] * yx + am1[j];
                float ub = (yx * yx + 1.0f) * pexp(am1[j] * plog(x / am1[j]) - yx * sq[j]);
                acc = (0.0f < x) && (u2[j] <= ub);
            }
            if (acc && !done[j]) { res[j] = beta[j] * x; done[j] = true; }
            all = all && done[j];
        }
        if (all) break;
    }
    std::memcpy(out, res, 32);
}

// Poisson: Packet4f (half draws), per-lane mean m. Lanes with m < 12 run the
// Knuth product loop in lockstep (one half draw per iteration); then the
// others run the tangent step with float constants (two half draws per
// iteration): psq = sqrt(2m), plm = log m, pg1 = m*plm - lgamma_approx(m+1).
void poissonv4(Vmt& e, const float mean[4], int32_t out[4]) {
    bool small[4], any_small = false, all_small = true;
    float nem[4], psq[4], plm[4], pg1[4], res[4] = {0, 0, 0, 0};
    for (int j = 0; j < 4; ++j) {
        small[j] = mean[j] < 12.0f;
        any_small |= small[j];
        all_small &= small[j];
        nem[j] = pexp(-mean[j]);
        psq[j] = std::sqrt(2.0f * mean[j]);
        plm[j] = plog(mean[j]);
        pg1[j] = mean[j] * plm[j] - plgamma_approx(mean[j] + 1.0f);
    }
    if (any_small) {
        float val[4] = {1, 1, 1, 1};
        int32_t cnt[4] = {0, 0, 0, 0};
        for (;;) {
            float u[4];
            ur4(e, u);
            bool any = false, c[4];
            for (int j = 0; j < 4; ++j) {
                val[j] = val[j] * u[j];
                c[j] = (nem[j] < val[j]) && small[j];
                any |= c[j];
            }
            if (!any) break;
            for (int j = 0; j < 4; ++j) cnt[j] += c[j] ? 1 : 0;
        }
        for (int j = 0; j < 4; ++j) res[j] = (float)cnt[j];
    }
    if (!all_small) {
        bool done[4];
        for (int j = 0; j < 4; ++j) done[j] = small[j];
        for (;;) {
            float u1[4], u2[4], fres[4], p12[4];
            bool c1[4];
            ur4(e, u1);
            for (int j = 0; j < 4; ++j) {
                float s, c;
                psincos(kPiF * u1[j], s, c);
                float yx = s / c;
                fres[j] = ptruncate(psq[j] * yx + mean[j]);
                float p1 = (yx * yx + 1.0f) * F(0.9);
                float p2 = pexp((fres[j] * plm[j] - plgamma_approx(fres[j] + 1.0f)) - pg1[j]);
                c1[j] = 0.0f <= fres[j];
                p12[j] = p1 * p2;
            }
            ur4(e, u2);
            bool all = true;
            for (int j = 0; j < 4; ++j) {
                bool acc = c1[j] && (u2[j] <= p12[j]);
                if (acc && !done[j]) { res[j] = fres[j]; done[j] = true; }
                all = all && done[j];
            }
            if (all) break;
        }
    }
    for (int j = 0; j < 4; ++j) out[j] = (int32_t)res[j];
}

}  // namespace

extern "C" {
void orc_fill_gamma_vop_f32(void* states, int64_t n_engines, float* out, int64_t per,
                            const float* alpha, const float* beta, int threads) {
    Vmt* e = (Vmt*)states;
    for_engines(n_engines, threads, [&](int64_t g) {
        int64_t b = g * per;
        for (int64_t i = 0; i < per; i += 8) gammav8(e[g], alpha + b + i, beta + b + i, out + b + i);
    });
}
void orc_fill_poisson_vop_i32(void* states, int64_t n_engines, int32_t* out, int64_t per,
                              const float* mean, int threads) {
    Vmt* e = (Vmt*)states;
    for_engines(n_engines, threads, [&](int64_t g) {
        int64_t b = g * per;
        for (int64_t i = 0; i < per; i += 4) poissonv4(e[g], mean + b + i, out + b + i);
    });
}
}

// ----------------------------------------------------------------------------
// DiscreteGen<int32_t, int32_t> and MultinomialGen (Discrete.h:23-207,
// 352-485; MvDists/Multinomial.h:25-134).
namespace {

// AliasMethod<int32_t, int32_t>::buildTable (Discrete.h:78-192), in double.
struct AliasTable {
    std::vector<int32_t> prob, alias;
    std::vector<float> fprob;  // AliasMethod<float, int32_t>: (float)f, or 1
    std::vector<double> dprob; // AliasMethod<double, int32_t>: f, or 1
    size_t bitsize = 0, bitmask = 0;
};

bool build_alias(const double* w, size_t msize, AliasTable& t) {
    double sum = 0;
    for (size_t i = 0; i < msize; ++i) sum += w[i];
    if (!std::isfinite(sum)) return false;
    const size_t nbsize = (size_t)std::ceil(std::log2((double)msize));
    const size_t psize = (size_t)1 << nbsize;
    t.prob.assign(psize, 0);
    t.fprob.assign(psize, 0.0f);
    t.dprob.assign(psize, 0.0);
    t.alias.assign(psize, 0);
    t.bitsize = nbsize;
    t.bitmask = ((size_t)(-1)) >> (sizeof(size_t) * 8 - nbsize);
    sum /= psize;
    std::vector<double> f(psize, 0.0);
    for (size_t i = 0; i < msize; ++i) f[i] = w[i] / sum;
    const int32_t imax = std::numeric_limits<int32_t>::max();
    size_t over = 0, under = 0, mm;
    while (over < psize && f[over] < 1) ++over;
    while (under < psize && f[under] >= 1) ++under;
    mm = under + 1;
    while (over < psize && under < psize) {
        t.prob[under] = (int32_t)(f[under] * (imax + 1.0));
        t.fprob[under] = (float)f[under];
        t.dprob[under] = f[under];
        t.alias[under] = (int32_t)over;
        f[over] += f[under] - 1;
        if (f[over] >= 1 || mm <= over) {
            for (under = mm; under < psize && f[under] >= 1; ++under);
            mm = under + 1;
        } else {
            under = over;
        }
        while (over < psize && f[over] < 1) ++over;
    }
    for (; over < psize; ++over) {
        if (f[over] >= 1) {
            t.prob[over] = imax;
            t.fprob[over] = 1.0f;
            t.dprob[over] = 1.0;
            t.alias[over] = (int32_t)over;
        }
    }
    if (under < psize) {
        t.prob[under] = imax;
        t.fprob[under] = 1.0f;
        t.dprob[under] = 1.0;
        t.alias[under] = (int32_t)under;
        for (under = mm; under < msize; ++under) {
            if (f[under] < 1) {
                t.prob[under] = imax;
                t.fprob[under] = 1.0f;
                t.dprob[under] = 1.0;
                t.alias[under] = (int32_t)under;
            }
        }
    }
    return true;
}

struct Discrete {
    std::vector<uint32_t> cdf;  // K < 16
    AliasTable at;              // K >= 16
    bool valid = false;         // AVX2: upper Packet4i of the last Packet8i
    int32_t cache[4];

    bool init(const double* w, size_t K) {
        if (K < 16) {
            std::vector<double> c;
            double acc = 0;
            for (size_t i = 0; i < K; ++i) c.push_back(acc += w[i]);
            for (double p : c) cdf.push_back((uint32_t)(p / c.back() * 0x80000000));
            return true;
        }
        return build_alias(w, K, at);
    }
    int32_t from_word_packet(uint32_t word) const {
        if (!cdf.empty()) {
            const int32_t rx = (int32_t)(word & 0x7FFFFFFFu);
            int32_t ret = (int32_t)cdf.size() - 1;
            for (size_t i = 0; i + 1 < cdf.size(); ++i) ret += (rx < (int32_t)cdf[i]) ? -1 : 0;
            return ret;
        }
        const int32_t albit = (int32_t)(word & (uint32_t)at.bitmask);
        const bool c = (int32_t)(word >> 1) < at.prob[albit];
        return c ? albit : at.alias[albit];
    }
    // packetOp<Packet4i> under AVX2 (Discrete.h:443-483).
    void packet4(Vmt& e, int32_t out[4]) {
        if (valid) {
            valid = false;
            std::memcpy(out, cache, 16);
            return;
        }
        uint64_t p[4];
        vmt_draw(e, p);
        int32_t ret[8];
        for (int j = 0; j < 8; ++j) ret[j] = from_word_packet((uint32_t)(p[j >> 1] >> (32 * (j & 1))));
        valid = true;
        std::memcpy(cache, ret + 4, 16);
        std::memcpy(out, ret, 16);
    }
    // operator() (Discrete.h:428-441): one full draw, first 32-bit word.
    int32_t scalar(Vmt& e) {
        uint64_t p[4];
        vmt_draw(e, p);
        const int32_t rx = (int32_t)(uint32_t)p[0];
        if (!cdf.empty()) {
            const int32_t r = rx & 0x7FFFFFFF;
            return (int32_t)(std::lower_bound(cdf.begin(), cdf.end() - 1, (uint32_t)r) - cdf.begin());
        }
        const size_t albit = (size_t)(int64_t)rx & at.bitmask;
        const uint32_t alx = (uint32_t)(rx >> 1);
        if (alx < (uint32_t)at.prob[albit]) return (int32_t)albit;
        return at.alias[albit];
    }
};

// MultinomialGen::generate(urng, samples) (Multinomial.h:60-133); out is
// K x samples column-major.
bool multinomial(Vmt& e, int32_t trials, const double* w, int64_t K, int64_t samples, int32_t* out) {
    std::vector<double> probs(w, w + K);
    Discrete disc;
    if (!disc.init(probs.data(), (size_t)K)) return false;
    double sum = 0;
    for (double v : probs) sum += v;
    for (double& v : probs) v /= sum;
    std::fill(out, out + K * samples, 0);
    if (trials < std::max<int64_t>(30 * (K - 1), 100)) {
        for (int64_t s = 0; s < samples; ++s) {
            int32_t* col = out + s * K;
            int32_t i = 0;
            for (; i < (trials & ~15); i += 16) {
                int32_t buf[16];
                for (int q = 0; q < 4; ++q) disc.packet4(e, buf + 4 * q);
                for (int j = 0; j < 16; ++j) col[buf[j]] += 1;
            }
            const int r = trials - i;
            int32_t buf[16];
            int j = 0;
            for (; j + 4 <= r; j += 4) disc.packet4(e, buf + j);
            for (; j < r; ++j) buf[j] = disc.scalar(e);
            for (j = 0; j < r; ++j) col[buf[j]] += 1;
        }
        return true;
    }
    std::vector<int32_t> r_trials((size_t)samples, trials), t((size_t)samples);
    double rest_p = 1 - probs[0];
    for (int64_t row = 0; row + 1 < K; ++row) {
        const float pp = row == 0 ? (float)probs[0] : (float)(probs[row] / rest_p);
        for (int64_t s = 0; s < samples; s += 8) {
            float pv[8];
            for (int q = 0; q < 8; ++q) pv[q] = pp;
            binomialv<8>(e, r_trials.data() + s, pv, t.data() + s);
        }
        for (int64_t s = 0; s < samples; ++s) {
            r_trials[(size_t)s] -= t[(size_t)s];
            out[s * K + row] = t[(size_t)s];
        }
        if (row > 0) rest_p -= probs[row];
    }
    for (int64_t s = 0; s < samples; ++s) out[s * K + K - 1] = r_trials[(size_t)s];
    return true;
}

}  // namespace

extern "C" {
// Alias table of AliasMethod<int32_t> for weights w (K >= 16); prob/alias
// must hold 2^ceil(log2 K) entries. Returns the bit size, or -1.
int orc_alias_table(const double* w, int64_t K, int32_t* prob, int32_t* alias) {
    AliasTable t;
    if (!build_alias(w, (size_t)K, t)) return -1;
    std::memcpy(prob, t.prob.data(), 4 * t.prob.size());
    std::memcpy(alias, t.alias.data(), 4 * t.alias.size());
    return (int)t.bitsize;
}

// AliasMethod<float, int32_t> (float branch of buildTable, Discrete.h:119-183).
int orc_alias_table_f64(const double* w, int64_t K, double* prob, int32_t* alias) {
    AliasTable t;
    if (!build_alias(w, (size_t)K, t)) return -1;
    std::memcpy(prob, t.dprob.data(), 8 * t.dprob.size());
    std::memcpy(alias, t.alias.data(), 4 * t.alias.size());
    return (int)t.bitsize;
}

// DiscreteGen<int32_t, double>::packetOp at Packet4i under AVX with Eigen 3.x
// (Discrete.h:699-779, the EIGEN_VECTORIZE_AVX branches). K < 16: one full
// draw gives 4 doubles rx (StdUniformRealGen<double>, bit_to_ur_double of the
// u64 lanes) and out = K - #{c in cdf : rx < c}, cdf_i = c_i / c_K in double.
// Otherwise idx = the 4 int32 words of a half() draw & bitmask, then one full
// draw's 4 doubles u: idx if u < prob[idx] (AliasMethod<double, int32_t>)
// else alias[idx].
void orc_fill_discrete_d_i32(void* states, int64_t n_engines, int32_t* out, int64_t per_engine, const double* cdf,
                             int32_t k, const double* prob, const int32_t* alias, uint32_t bitmask) {
    Vmt* e = (Vmt*)states;
    for (int64_t g = 0; g < n_engines; ++g) {
        int32_t* o = out + g * per_engine;
        for (int64_t i = 0; i < per_engine; i += 4) {
            uint64_t p[4];
            if (cdf) {
                vmt_draw(e[g], p);
                for (int j = 0; j < 4; ++j) {
                    const double rx = bit_to_ur_double(p[j]);
                    int32_t r = k;
                    for (int c = 0; c < k; ++c) r -= rx < cdf[c] ? 1 : 0;
                    o[i + j] = r;
                }
            } else {
                uint64_t h[2];
                vmt_half(e[g], h);
                uint32_t idx[4];
                for (int j = 0; j < 4; ++j) idx[j] = (uint32_t)(h[j >> 1] >> (32 * (j & 1))) & bitmask;
                vmt_draw(e[g], p);
                for (int j = 0; j < 4; ++j)
                    o[i + j] = bit_to_ur_double(p[j]) < prob[idx[j]] ? (int32_t)idx[j] : alias[idx[j]];
            }
        }
    }
}

int orc_alias_table_f32(const double* w, int64_t K, float* prob, int32_t* alias) {
    AliasTable t;
    if (!build_alias(w, (size_t)K, t)) return -1;
    std::memcpy(prob, t.fprob.data(), 4 * t.fprob.size());
    std::memcpy(alias, t.alias.data(), 4 * t.alias.size());
    return (int)t.bitsize;
}

// Batched MultinomialGen: instance i on engine i; out_i K_i x samples_i.
int orc_multinomial(void* states, int64_t n_inst, const int32_t* trials, const int64_t* K,
                    const int64_t* samples, const double* w, int32_t* out) {
    Vmt* e = (Vmt*)states;
    int64_t woff = 0, ooff = 0;
    for (int64_t i = 0; i < n_inst; ++i) {
        if (!multinomial(e[i], trials[i], w + woff, K[i], samples[i], out + ooff)) return -1;
        woff += K[i];
        ooff += K[i] * samples[i];
    }
    return 0;
}

// DiscreteGen<int32_t, int32_t> fill (packet path, AVX2 Packet4i halves of a
// Packet8i): per engine `per` draws, a multiple of 4.
int orc_fill_discrete_i32(void* states, int64_t n_engines, int32_t* out, int64_t per, const double* w,
                          int64_t K) {
    Vmt* e = (Vmt*)states;
    for (int64_t g = 0; g < n_engines; ++g) {
        Discrete d;
        if (!d.init(w, (size_t)K)) return -1;
        for (int64_t i = 0; i < per; i += 4) d.packet4(e[g], out + g * per + i);
    }
    return 0;
}
}

// The reference benchmark's own comparator (benchmark/benchmark.cpp:531-543):
// scalar std::mt19937_64 and std::normal_distribution<float>. It is not part
// of the restatement and no check uses it; it is timed beside the restatement
// as a CPU baseline. Thread t fills the t-th contiguous slice from its own
// std::mt19937_64(seed + t).
extern "C" void orc_std_normal_f32(float* out, int64_t n, uint64_t seed, int threads) {
    threads = std::max(1, threads);
    std::vector<std::thread> pool;
    const int64_t chunk = (n + threads - 1) / threads;
    for (int t = 0; t < threads; ++t) {
        const int64_t lo = std::min<int64_t>(n, t * chunk), hi = std::min<int64_t>(n, lo + chunk);
        pool.emplace_back([=] {
            std::mt19937_64 rng(seed + (uint64_t)t);
            std::normal_distribution<float> nd;
            for (int64_t i = lo; i < hi; ++i) out[i] = nd(rng);
        });
    }
    for (auto& th : pool) th.join();
}


// ------------------------------------------------ scalar-engine (P8) feeds ---
// P8_mt19937_64 (ParallelRandomEngineAdaptor<uint64_t, Vmt19937_64, 8>,
// PacketRandomEngine.h:429-517) as a scalar full-bit engine: operator()
// returns buf[cnt++], a refill drawing Vmt (seed) then Vmt (seed + 4). Packet
// code takes it through RawbitsMaker<..., scalar_fullbit> with the argument
// order of Clang (left to right):
//  mode 0, raw: operator() in order, four scalars per draw;
//  mode 1, Packet8f rawbits_34 (arch/AVX/RandUtils.h:75-131):
//          _mm256_setr_epi64x(a, b, c, 0), the 8x32 permutation (0,1,2,7,3,
//          4,5,7) and the byte shuffle (0..11, 3, 7, 11, 11) per 128 bits;
//  mode 2, Packet4i rawbits (arch/SSE/RandUtils.h:24-41):
//          _mm_set_epi64x(first, second), lane 0 = the second scalar.
namespace {
struct P8Scalar {
    Vmt* v0;
    Vmt* v1;
    uint64_t buf[8];
    int cnt;   // 8: empty
    int mode;
    uint64_t next() {
        if (cnt >= 8) {
            vmt_draw(*v0, buf);
            vmt_draw(*v1, buf + 4);
            cnt = 0;
        }
        return buf[cnt++];
    }
};

void p8s_full(P8Scalar* f, uint64_t out[4]) {
    if (f->mode == 0) {
        for (int l = 0; l < 4; ++l) out[l] = f->next();
        return;
    }
    // rawbits_34: bytes of the 8 32-bit words after set/permute/shuffle
    const uint64_t a = f->next(), b = f->next(), c = f->next();
    uint32_t w[8] = {(uint32_t)a, (uint32_t)(a >> 32), (uint32_t)b, 0,
                     (uint32_t)(b >> 32), (uint32_t)c, (uint32_t)(c >> 32), 0};
    auto top = [](uint32_t x) { return x >> 24; };
    w[3] = top(w[0]) | (top(w[1]) << 8) | (top(w[2]) << 16) | (top(w[2]) << 24);
    w[7] = top(w[4]) | (top(w[5]) << 8) | (top(w[6]) << 16) | (top(w[6]) << 24);
    for (int l = 0; l < 4; ++l) out[l] = (uint64_t)w[2 * l] | ((uint64_t)w[2 * l + 1] << 32);
}

void p8s_half(P8Scalar* f, uint64_t out[2]) {
    const uint64_t first = f->next(), second = f->next();
    out[0] = second;
    out[1] = first;
}

}  // namespace

// dist (erd_dist codes) filled per adaptor with P8 adaptor k (records 2k, 2k+1)
// as its engine; mode_override >= 0 forces a packing (0 for operator()).
// Record 2k's `reserved` keeps the buffer position (scalars of the last
// refill returned; its draws are the records' last ones).
extern "C" int orc_p8_fill(void* states, int64_t n_p8, int dist, int mode_override, void* out, int64_t per,
                           double a, double b) {
    Vmt* st = (Vmt*)states;
    for (int64_t k = 0; k < n_p8; ++k) {
        P8Scalar f;
        f.v0 = &st[2 * k];
        f.v1 = &st[2 * k + 1];
        const int b0 = (int)f.v0->reserved;
        f.cnt = 8;
        if (b0 > 0) {
            for (int l = 0; l < 4; ++l) {
                f.buf[l] = temper(f.v0->mt[(f.v0->idx - 1) * 4 + l]);
                f.buf[4 + l] = temper(f.v1->mt[(f.v1->idx - 1) * 4 + l]);
            }
            f.cnt = b0;
        }
        const bool ints = dist == 10 || dist == 11;
        f.mode = mode_override >= 0 ? mode_override : dist == 0 ? 0 : ints ? 2 : 1;
        Vmt v;
        std::memset(&v, 0, sizeof(v));
        v.half_valid = -1;
        v.reserved = (uint64_t)(uintptr_t)&f;
        float* of = (float*)out + k * per;
        int32_t* oi = (int32_t*)out + k * per;
        switch (dist) {
            case 0: orc_fill_u64(&v, 1, (uint64_t*)out + k * per, per / 4, 1); break;
            case 1: orc_fill_uniform_f32(&v, 1, of, per, (float)a, (float)b, !(a == 0.0 && b == 1.0), 1); break;
            case 3: orc_fill_normal_f32(&v, 1, of, per, 0.f, 1.f, 0, 1); break;
            case 4: orc_fill_normal_f32(&v, 1, of, per, (float)a, (float)b, 1, 1); break;
            case 6: orc_fill_exponential_f32(&v, 1, of, per, (float)a, 1); break;
            case 7: orc_fill_cauchy_f32(&v, 1, of, per, (float)a, (float)b, 1); break;
            case 8: orc_fill_studentt_f32(&v, 1, of, per, (float)a, 1); break;
            case 9: orc_fill_gamma_f32(&v, 1, of, per, (float)a, (float)b, 1); break;
            case 10: orc_fill_poisson_i32(&v, 1, oi, per, a, 1); break;
            case 11: orc_fill_binomial_i32(&v, 1, oi, per, (int32_t)a, b, 1); break;
            case 16: orc_fill_lognormal_f32(&v, 1, of, per, (float)a, (float)b, 1); break;
            case 17: orc_fill_weibull_f32(&v, 1, of, per, (float)a, (float)b, 1); break;
            case 18: orc_fill_extreme_value_f32(&v, 1, of, per, (float)a, (float)b, 1); break;
            default: return -1;
        }
        f.v0->reserved = f.cnt >= 8 ? 0 : (uint64_t)f.cnt;
    }
    return 0;
}

// ------------------------------------------------------------ lane layouts ---
// Fills at the SSE2/NEON (W = 4) and AVX-512 (W = 16) layouts. W = 4: the
// 2-lane Vmt19937_64 = Pmt19937_64<Packet4i> (PacketRandomEngine.h:587-588)
// seeded x0 has lanes x0, x0+1, which are lanes 0-1 of the 4-lane engine
// seeded x0; one call makes a Packet4f / Packet2d (arch/SSE/RandUtils.h:
// 70-115). W = 16: a Packet16f / Packet8d is two calls of the AVX2 engine,
// the first in the low half (arch/AVX512/RandUtils.h:37-46, arguments
// evaluated left to right). The packet ops are those of the AVX2 fills at
// the wider or narrower packet.
namespace {
// One raw packet at layout W: its u64 words (2, 4 or 8).
inline int raw_packet(Vmt& e, int W, uint64_t* p) {
    if (W == 4) {
        uint64_t d[4];
        vmt_draw(e, d);
        p[0] = d[0];
        p[1] = d[1];
        return 2;
    }
    vmt_draw(e, p);
    if (W == 16) {
        vmt_draw(e, p + 4);
        return 8;
    }
    return 4;
}
// Packet of W floats: float j = 32-bit word j of the packet, low word first.
inline void ur_packet(Vmt& e, int W, float* u) {
    uint64_t p[8];
    raw_packet(e, W, p);
    for (int j = 0; j < W; ++j) u[j] = bit_to_ur_float((uint32_t)(p[j >> 1] >> (32 * (j & 1))));
}
}  // namespace

extern "C" {

// Raw draws at layout W: calls of 2 u64 (W = 4) or 4 u64 (W = 8, 16).
void orc_fill_u64_l(void* states, int64_t n_engines, uint64_t* out, int64_t per, int W) {
    Vmt* e = (Vmt*)states;
    const int width = W == 4 ? 2 : 4;
    for (int64_t g = 0; g < n_engines; ++g)
        for (int64_t i = 0; i < per; i += width) raw_packet(e[g], W == 4 ? 4 : 8, out + g * per + i);
}

// Float fills at layout W (per a multiple of W). kind: 0 uniformReal (mode 1:
// a + u*(b-a)), 1 normal (StdNormalGen, sine packet then the cached cosine
// packet; mode 0 standard, 1 z*b + a, 3 exp(z*b + a)), 2 exponential(a),
// 3 Cauchy(a, b), 4 Weibull(a, b), 5 extreme value(a, b), 6 Student-t(a).
void orc_fill_f32_l(void* states, int64_t n_engines, int W, int kind, float* out, int64_t per, float a, float b,
                    int mode) {
    Vmt* e = (Vmt*)states;
    const float slope = b - a, inv_a = 1 / a, k = -2 / a;
    for (int64_t g = 0; g < n_engines; ++g) {
        float* o = out + g * per;
        float u1[16], u2[16], cache[16];
        bool valid = false;
        for (int64_t i = 0; i < per; i += W) {
            float* v = o + i;
            switch (kind) {
                case 0:
                    ur_packet(e[g], W, u1);
                    for (int j = 0; j < W; ++j) v[j] = mode ? u1[j] * slope + a : u1[j];
                    break;
                case 1:
                    if (valid) {
                        valid = false;
                        for (int j = 0; j < W; ++j) v[j] = cache[j];
                    } else {
                        valid = true;
                        ur_packet(e[g], W, u1);
                        ur_packet(e[g], W, u2);
                        for (int j = 0; j < W; ++j) {
                            const float radius = std::sqrt(-2.0f * plog(1.0f - u1[j]));
                            float sn, cs;
                            psincos(F(2 * kPi) * u2[j], sn, cs);
                            cache[j] = radius * cs;
                            v[j] = radius * sn;
                        }
                    }
                    if (mode == 1) for (int j = 0; j < W; ++j) v[j] = v[j] * b + a;
                    if (mode == 3) for (int j = 0; j < W; ++j) v[j] = pexp(v[j] * b + a);
                    break;
                case 2:
                    ur_packet(e[g], W, u1);
                    for (int j = 0; j < W; ++j) v[j] = -(plog(1.0f - u1[j]) / a);
                    break;
                case 3:
                    ur_packet(e[g], W, u1);
                    for (int j = 0; j < W; ++j) {
                        float sn, cs;
                        psincos(kPiF * (u1[j] - F(0.5)), sn, cs);
                        v[j] = a + b * (sn / cs);
                    }
                    break;
                case 4:
                    ur_packet(e[g], W, u1);
                    for (int j = 0; j < W; ++j) v[j] = pexp(plog(-plog(1.0f - u1[j])) * inv_a) * b;
                    break;
                case 5:
                    ur_packet(e[g], W, u1);
                    for (int j = 0; j < W; ++j) v[j] = a - plog(-plog(u1[j] + 1.1920928955078125e-07f / 8)) * b;
                    break;
                case 6:
                    ur_packet(e[g], W, u1);
                    ur_packet(e[g], W, u2);
                    for (int j = 0; j < W; ++j) {
                        const float radius = std::sqrt(a * (pexp(plog(1.0f - u1[j]) * k) - 1.0f));
                        float sn, cs;
                        psincos(F(2 * kPi) * u2[j], sn, cs);
                        v[j] = radius * sn;
                    }
                    break;
            }
        }
    }
}

// Double fills at layout W (per a multiple of W/2 doubles). kind: 0
// uniformReal (mode 1: a + u*(b-a)), 1 normal (mode 1: z*b + a), 2
// exponential(a), 3 Cauchy(a, b).
void orc_fill_f64_l(void* states, int64_t n_engines, int W, int kind, double* out, int64_t per, double a,
                    double b, int mode) {
    Vmt* e = (Vmt*)states;
    const int PD = W / 2;
    for (int64_t g = 0; g < n_engines; ++g) {
        double* o = out + g * per;
        uint64_t p1[8], p2[8];
        double cache[8];
        bool valid = false;
        for (int64_t i = 0; i < per; i += PD) {
            double* v = o + i;
            switch (kind) {
                case 0:
                    raw_packet(e[g], W, p1);
                    for (int j = 0; j < PD; ++j) {
                        const double u = bit_to_ur_double(p1[j]);
                        v[j] = mode ? u * (b - a) + a : u;
                    }
                    break;
                case 1:
                    if (valid) {
                        valid = false;
                        for (int j = 0; j < PD; ++j) v[j] = cache[j];
                    } else {
                        valid = true;
                        raw_packet(e[g], W, p1);
                        raw_packet(e[g], W, p2);
                        for (int j = 0; j < PD; ++j) {
                            const double radius = std::sqrt(-2.0 * plog_d(1.0 - bit_to_ur_double(p1[j])));
                            double sn, cs;
                            psincos_d((2 * kPi) * bit_to_ur_double(p2[j]), sn, cs);
                            cache[j] = radius * cs;
                            v[j] = radius * sn;
                        }
                    }
                    if (mode == 1) for (int j = 0; j < PD; ++j) v[j] = v[j] * b + a;
                    break;
                case 2:
                    raw_packet(e[g], W, p1);
                    for (int j = 0; j < PD; ++j) v[j] = -(plog_d(1.0 - bit_to_ur_double(p1[j])) / a);
                    break;
                case 3:
                    raw_packet(e[g], W, p1);
                    for (int j = 0; j < PD; ++j) {
                        double sn, cs;
                        psincos_d(kPi * (bit_to_ur_double(p1[j]) - 0.5), sn, cs);
                        v[j] = a + b * (sn / cs);
                    }
                    break;
            }
        }
    }
}

}  // extern "C"
