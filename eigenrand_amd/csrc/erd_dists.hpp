// Per-distribution arithmetic shared by the fill kernels and the segment
// program kernel: fixed-consumption unit functors (op_fixed) and rejection
// candidate policies (op_reject). Each follows the reference's packetOp line
// by line in the unfused float order.
#pragma once
#include "erd_device.hpp"

namespace erd {

// ============================================================ fixed units ===
// Each functor: CALLS full draws per unit; units(per) = units for `per`
// outputs; emit(per, u, l, w) writes the outputs of unit u from the tempered
// words w[0..CALLS) of u64 lane l; rebase(off) moves its pointers to the block
// of one engine.

struct RawU64 {  // out[4u + l] = lane l of call u (PacketRandomEngine.h:295-310)
    static constexpr int CALLS = 1;
    uint64_t* out;
    __device__ void rebase(int64_t off) { out += off; }
    __device__ static int64_t units(int64_t n) { return n / 4; }
    __device__ uint64_t value(uint64_t w) const { return w; }
    __device__ void emit(int64_t per, int64_t u, int l, const uint64_t* w) const {
        out[4 * u + l] = w[0];
    }
};

struct UniformF32 {  // Basic.h:584-590 / 626-634; float j of a call = word j
    static constexpr int CALLS = 1;
    float* out;
    float a, slope;
    int affine;
    __device__ void rebase(int64_t off) { out += off; }
    __device__ static int64_t units(int64_t n) { return n / 8; }
    __device__ float value(uint32_t x) const {
        const float v = ur_f32(x);
        return affine ? v * slope + a : v;
    }
    __device__ void emit(int64_t per, int64_t u, int l, const uint64_t* w) const {
        *reinterpret_cast<float2*>(out + 8 * u + 2 * l) = make_float2(value(lo32(w[0])), value(hi32(w[0])));
    }
    __device__ static int64_t full_units(int64_t n) { return n / 8; }
    // floats 4k .. 4k+3 of the call (lanes 2k, 2k+1) in one 16-byte store
    __device__ void emit2(int64_t per, int64_t u, int k, const uint64_t (&w)[1][2], bool) const {
        float v[4] = {ur_f32(lo32(w[0][0])), ur_f32(hi32(w[0][0])), ur_f32(lo32(w[0][1])), ur_f32(hi32(w[0][1]))};
        if (affine) {
#pragma unroll
            for (int i = 0; i < 4; ++i) v[i] = v[i] * slope + a;
        }
        *reinterpret_cast<float4*>(out + 8 * u + 4 * k) = make_float4(v[0], v[1], v[2], v[3]);
    }
};

struct UniformF64 {  // arch/AVX/RandUtils.h:212-225: 4 doubles per call
    static constexpr int CALLS = 1;
    double* out;
    double a, slope;
    int affine;
    __device__ void rebase(int64_t off) { out += off; }
    __device__ static int64_t units(int64_t n) { return n / 4; }
    __device__ double value(uint64_t w) const {
        const double v = ur_f64(w);
        return affine ? v * slope + a : v;
    }
    __device__ void emit(int64_t per, int64_t u, int l, const uint64_t* w) const {
        out[4 * u + l] = value(w[0]);
    }
};

// StdNormalGen / NormalGen at Packet4d (NormalExp.h:60-82, 127-135): draws
// 2m and 2m+1 give u1 and u2 of four Box-Muller pairs; out[8m + l] = r sin and
// out[8m + 4 + l] = r cos (the packet the generator caches for its next call).
// MODE 0 standard, 1 = z * sd + mean.
template <int MODE>
struct NormalF64 {
    static constexpr int CALLS = 2;
    double* out;
    double mean, sd;
    __device__ void rebase(int64_t off) { out += off; }
    __device__ static int64_t units(int64_t n) { return (n + 7) / 8; }
    __device__ void emit(int64_t per, int64_t u, int l, const uint64_t* w) const {
        const double u1 = 1.0 - ur_f64(w[0]);  // in [2^-52, 1], exact
        const double u2 = ur_f64(w[1]);
        const double radius = sqrt(-2.0 * plog_pos_f64(u1));
        double s, c;
        psincos_f64((2.0 * kPi) * u2, s, c);
        double sn = radius * s, cs = radius * c;
        if (MODE == 1) {
            sn = sn * sd + mean;
            cs = cs * sd + mean;
        }
        out[8 * u + l] = sn;
        if (8 * u + 8 <= per) out[8 * u + 4 + l] = cs;
    }
};

struct ExponentialF64 {  // NormalExp.h:391-398 at Packet4d
    static constexpr int CALLS = 1;
    double* out;
    double lambda;
    __device__ void rebase(int64_t off) { out += off; }
    __device__ static int64_t units(int64_t n) { return n / 4; }
    __device__ double value(uint64_t w) const { return -(plog_pos_f64(1.0 - ur_f64(w)) / lambda); }
    __device__ void emit(int64_t per, int64_t u, int l, const uint64_t* w) const {
        out[4 * u + l] = value(w[0]);
    }
};

struct CauchyF64 {  // NormalExp.h:843-855 at Packet4d
    static constexpr int CALLS = 1;
    double* out;
    double a, b;
    __device__ void rebase(int64_t off) { out += off; }
    __device__ static int64_t units(int64_t n) { return n / 4; }
    __device__ double value(uint64_t w) const {
        double s, c;
        psincos_f64(kPi * (ur_f64(w) - 0.5), s, c);
        return a + b * (s / c);
    }
    __device__ void emit(int64_t per, int64_t u, int l, const uint64_t* w) const {
        out[4 * u + l] = value(w[0]);
    }
};

// P8_mt19937_64 / P8_mt19937_64_32 (ParallelRandomEngineAdaptor<UInt,
// Vmt19937_64, 8>, PacketRandomEngine.h:429-517): one adaptor = the Vmt pair
// (seed, seed+4); a refill writes draw k of Vmt i to buffer words
// [8i, 8i+8) in u64 (MODE 0), at [16k + 8i, 16k + 8i + 8) as 32-bit words
// low half first (MODE 1), or as bit_to_ur_float of those words
// (uniform_real(), MODE 2). Engine g of the set is Vmt (g & 1) of adaptor g/2.
template <int MODE>
struct P8Fill {
    static constexpr int CALLS = 1;
    void* out;
    int64_t per_v;  // outputs per Vmt engine (half of the adaptor's block)
    int half;
    __device__ void rebase(int64_t off) {
        const int64_t g = off / per_v;
        half = (int)(g & 1);
        const int64_t base = (g >> 1) * 2 * per_v;
        if (MODE == 0) out = (uint64_t*)out + base;
        else out = (uint32_t*)out + base;
    }
    __device__ static int64_t units(int64_t n) { return MODE == 0 ? n / 4 : n / 8; }
    __device__ void emit(int64_t per, int64_t u, int l, const uint64_t* w) const {
        if (MODE == 0) {
            ((uint64_t*)out)[8 * u + 4 * half + l] = w[0];
        } else if (MODE == 1) {
            *reinterpret_cast<uint2*>((uint32_t*)out + 16 * u + 8 * half + 2 * l) =
                make_uint2(lo32(w[0]), hi32(w[0]));
        } else {
            *reinterpret_cast<float2*>((float*)out + 16 * u + 8 * half + 2 * l) =
                make_float2(ur_f32(lo32(w[0])), ur_f32(hi32(w[0])));
        }
    }
};

// Box-Muller pair of StdNormalGen::packetOp (NormalExp.h:60-82): calls A, B
// give u1 = 1 - U(A), u2 = U(B); the first packet is r*sin, the cached second
// packet r*cos. A fill of 8 mod 16 outputs ends with an unused cos packet.
template <int MODE>  // 0 = StdNormalGen, 1 = NormalGen(mean, sd), 2 = NormalVGen,
                     // 3 = LognormalGen(mean, sd) = pexp of NormalGen (NormalExp.h:171-210),
                     // 4 = LognormalVGen = pexp of NormalVGen (NormalExp.h:229-233)
struct NormalF32 {
    static constexpr int CALLS = 2;
    float* out;
    float mean, sd;
    const float* vmean;
    const float* vsd;
    GenState* gs;  // reused generator: receives the cos packet of a trailing half unit
    __device__ void rebase(int64_t off) {
        out += off;
        if (MODE == 2 || MODE == 4) {  // NULL: the scalar mean / sd for every element
            if (vmean) vmean += off;
            if (vsd) vsd += off;
        }
    }
    __device__ float vm(int64_t i) const { return vmean ? vmean[i] : mean; }
    __device__ float vs(int64_t i) const { return vsd ? vsd[i] : sd; }
    __device__ static int64_t units(int64_t n) { return (n + 15) / 16; }
    __device__ void emit(int64_t per, int64_t u, int l, const uint64_t* w) const {
        float sn[2], cs[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const float u1 = one_minus_ur_f32(h ? hi32(w[0]) : lo32(w[0]));
            const float u2 = ur_f32(h ? hi32(w[1]) : lo32(w[1]));
            // -2 ln(u1) for u1 in [2^-23, 1] is +-0 or >= 2^-22
            const float radius = sqrt_cr_normal(-2.0f * plog_pos(u1));
            float s, c;
            psincos_pos(ERD_F(2 * kPi) * u2, s, c);
            sn[h] = radius * s;
            cs[h] = radius * c;
        }
        const int64_t base = 16 * u + 2 * l;
        const bool full = 16 * u + 16 <= per;
        if (!full && gs) {  // the generator keeps the raw cos packet for its next call
            gs->v[2 * l] = cs[0];
            gs->v[2 * l + 1] = cs[1];
        }
        if (MODE == 1) {
#pragma unroll
            for (int h = 0; h < 2; ++h) { sn[h] = sn[h] * sd + mean; cs[h] = cs[h] * sd + mean; }
        } else if (MODE == 3) {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                sn[h] = pexp(sn[h] * sd + mean);
                cs[h] = pexp(cs[h] * sd + mean);
            }
        } else if (MODE == 2 || MODE == 4) {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                sn[h] = sn[h] * vs(base + h) + vm(base + h);
                if (full) cs[h] = cs[h] * vs(base + 8 + h) + vm(base + 8 + h);
                if (MODE == 4) {
                    sn[h] = pexp(sn[h]);
                    if (full) cs[h] = pexp(cs[h]);
                }
            }
        }
        *reinterpret_cast<float2*>(out + base) = make_float2(sn[0], sn[1]);
        if (full) *reinterpret_cast<float2*>(out + base + 8) = make_float2(cs[0], cs[1]);
    }
};

// Box-Muller with the hardware transcendentals (the fast precision mode):
// the same engine words, the same pairing and the same output order as
// NormalF32 (float j of draw 2m gives u1 = 1 - U, float j of draw 2m+1 gives
// u2; out[16m + j] = r sin, out[16m + 8 + j] = r cos), but
//   r = v_sqrt(-2 ln2 * v_log(u1))           (v_log is log2)
//   sin, cos = v_sin, v_cos of 1 + u2        (they take revolutions and have
//                                             period 1, so 1 + u2, which is
//                                             exact, stands for 2 pi u2)
// and four outputs per 16-byte store. The stated error bound against the
// bit-exact mode is asserted by the tests.
template <int AFFINE>  // 1: z * sd + mean
struct NormalF32Fast {
    static constexpr int CALLS = 2;
    float* out;
    float mean, sd;
    __device__ void rebase(int64_t off) { out += off; }
    __device__ static int64_t units(int64_t n) { return (n + 15) / 16; }
    __device__ static int64_t full_units(int64_t n) { return n / 16; }
    __device__ void emit2(int64_t per, int64_t u, int k, const uint64_t (&w)[2][2], bool full) const {
        float sn[4], cs[4];
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const uint32_t a = h ? hi32(w[0][j]) : lo32(w[0][j]);
                const uint32_t b = h ? hi32(w[1][j]) : lo32(w[1][j]);
                const float u1 = one_minus_ur_f32(a);
                const float rev = as_f((b & 0x7FFFFFu) | 0x3F800000u);
                const float r = __builtin_amdgcn_sqrtf(-1.38629436111989061f * __builtin_amdgcn_logf(u1));
                sn[2 * j + h] = r * __builtin_amdgcn_sinf(rev);
                cs[2 * j + h] = r * __builtin_amdgcn_cosf(rev);
            }
        if (AFFINE) {
#pragma unroll
            for (int i = 0; i < 4; ++i) { sn[i] = sn[i] * sd + mean; cs[i] = cs[i] * sd + mean; }
        }
        float* o = out + 16 * u + 4 * k;
        *reinterpret_cast<float4*>(o) = make_float4(sn[0], sn[1], sn[2], sn[3]);
        if (full) *reinterpret_cast<float4*>(o + 8) = make_float4(cs[0], cs[1], cs[2], cs[3]);
    }
};

struct WeibullF32 {  // NormalExp.h:649-655: b * exp(log(-log(1 - U)) * (1/a))
    static constexpr int CALLS = 1;
    float* out;
    float inv_a, b;  // 1/a rounded in float, as the reference's pset1(1 / a)
    __device__ void rebase(int64_t off) { out += off; }
    __device__ static int64_t units(int64_t n) { return n / 8; }
    __device__ float value(uint32_t w) const {
        const float x = 1.0f - ur_f32(w);
        return pexp(plog(-plog(x)) * inv_a) * b;
    }
    __device__ void emit(int64_t per, int64_t u, int l, const uint64_t* w) const {
        *reinterpret_cast<float2*>(out + 8 * u + 2 * l) = make_float2(value(lo32(w[0])), value(hi32(w[0])));
    }
};

struct ExtremeValueF32 {  // NormalExp.h:725-731: a - log(-log(U + eps/8)) * b
    static constexpr int CALLS = 1;
    float* out;
    float a, b;
    __device__ void rebase(int64_t off) { out += off; }
    __device__ static int64_t units(int64_t n) { return n / 8; }
    __device__ float value(uint32_t w) const { return a - plog(-plog(nzur_f32(w))) * b; }
    __device__ void emit(int64_t per, int64_t u, int l, const uint64_t* w) const {
        *reinterpret_cast<float2*>(out + 8 * u + 2 * l) = make_float2(value(lo32(w[0])), value(hi32(w[0])));
    }
};

struct ExponentialF32 {  // NormalExp.h:391-398
    static constexpr int CALLS = 1;
    float* out;
    float lambda;
    float scale;  // 1 for ExponentialGen; beta for GammaGen(alpha == 1) (NormalExp.h:561-565)
    int gamma;
    __device__ void rebase(int64_t off) { out += off; }
    __device__ static int64_t units(int64_t n) { return n / 8; }
    __device__ float value(uint32_t w) const {
        const float x = -(plog_pos(1.0f - ur_f32(w)) / lambda);
        return gamma ? scale * x : x;
    }
    __device__ void emit(int64_t per, int64_t u, int l, const uint64_t* w) const {
        *reinterpret_cast<float2*>(out + 8 * u + 2 * l) = make_float2(value(lo32(w[0])), value(hi32(w[0])));
    }
};

struct CauchyF32 {  // NormalExp.h:843-855
    static constexpr int CALLS = 1;
    float* out;
    float a, b;
    __device__ void rebase(int64_t off) { out += off; }
    __device__ static int64_t units(int64_t n) { return n / 8; }
    __device__ float value(uint32_t w) const {
        float s, c;
        psincos(ERD_F(kPi) * (ur_f32(w) - 0.5f), s, c);
        return a + b * (s / c);
    }
    __device__ void emit(int64_t per, int64_t u, int l, const uint64_t* w) const {
        *reinterpret_cast<float2*>(out + 8 * u + 2 * l) = make_float2(value(lo32(w[0])), value(hi32(w[0])));
    }
};

struct StudentTF32 {  // NormalExp.h:285-301 (sine from the reference's psincos)
    static constexpr int CALLS = 2;
    float* out;
    float n, k;  // k = -2/n in float
    __device__ void rebase(int64_t off) { out += off; }
    __device__ static int64_t units(int64_t n_) { return n_ / 8; }
    __device__ void emit(int64_t per, int64_t u, int l, const uint64_t* w) const {
        float v[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const float u1 = 1.0f - ur_f32(h ? hi32(w[0]) : lo32(w[0]));
            const float u2 = ur_f32(h ? hi32(w[1]) : lo32(w[1]));
            const float radius = sqrtf(n * (pexp(plog_pos(u1) * k) - 1.0f));
            float s, c;
            psincos_pos(ERD_F(2 * kPi) * u2, s, c);
            v[h] = radius * s;
        }
        *reinterpret_cast<float2*>(out + 8 * u + 2 * l) = make_float2(v[0], v[1]);
    }
};

// ------------------------------------------------------------ lane layouts ---
// The reference's packet width follows the ISA it is built for
// (PacketRandomEngine.h:585-588; arch/SSE/RandUtils.h:70-115;
// arch/AVX512/RandUtils.h:37-46). Besides the AVX2 layout above (W = 8
// floats per packet, one 4-lane call), two more are built:
//  - W = 4, SSE2 / NEON: Vmt19937_64 = Pmt19937_64<Packet4i>, a 2-lane engine
//    seeded x0, x0+1 = lanes 0-1 of the 4-lane record seeded x0. A call is
//    u64 lanes 0-1 of a draw (lanes 2-3 are stepped and never read): 4
//    floats, 2 doubles or 2 raw u64 per packet.
//  - W = 16, AVX-512: the AVX2 engine; a Packet16f / Packet8d is two calls,
//    the first in the low half (`_mm512_inserti64x4(cast(rng()), rng(), 1)`
//    with its arguments evaluated left to right, as the reference's default
//    Clang preset does). Element-wise fills therefore equal the AVX2 ones;
//    only samplers that pair two packets (Box-Muller, Student-t) differ.
// Items read u64 lanes 0..LANES-1 of each call (op_fixed).
template <int W>
struct Layout {
    static constexpr int LANES = W == 4 ? 2 : 4;  // u64 lanes read per call
    static constexpr int FPC = 2 * LANES;         // floats per call
    static constexpr int PC = W / FPC;            // calls per float packet
};

// Element-wise float distribution E (E::value of a 32-bit word) at W = 4:
// floats 4k..4k+3 of a block are the words of call k.
template <class E>
struct ElemF32W4 {
    static constexpr int CALLS = 1, LANES = 2;
    E d;
    __device__ void rebase(int64_t off) { d.out += off; }
    __device__ static int64_t units(int64_t n) { return n / 4; }
    __device__ void emit(int64_t per, int64_t u, int l, const uint64_t* w) const {
        *reinterpret_cast<float2*>(d.out + 4 * u + 2 * l) = make_float2(d.value(lo32(w[0])), d.value(hi32(w[0])));
    }
};

// Element-wise 64-bit distribution E (raw u64 or a double per u64 lane) at
// W = 4: elements 2k, 2k+1 of a block are lanes 0-1 of call k.
template <class E>
struct ElemU64W4 {
    static constexpr int CALLS = 1, LANES = 2;
    E d;
    __device__ void rebase(int64_t off) { d.out += off; }
    __device__ static int64_t units(int64_t n) { return n / 2; }
    __device__ void emit(int64_t per, int64_t u, int l, const uint64_t* w) const {
        d.out[2 * u + l] = d.value(w[0]);
    }
};

// Box-Muller pair of the exact normal fills: words a (u1 = 1 - U) and b (u2).
__device__ __forceinline__ void box_muller_f32(uint32_t a, uint32_t b, float& sn, float& cs) {
    const float u1 = one_minus_ur_f32(a);
    const float u2 = ur_f32(b);
    const float radius = sqrt_cr_normal(-2.0f * plog_pos(u1));
    float s, c;
    psincos_pos(ERD_F(2 * kPi) * u2, s, c);
    sn = radius * s;
    cs = radius * c;
}

// StdNormalGen / NormalGen / LognormalGen (NormalExp.h:60-82, 127-135,
// 197-210) at W = 4 or 16: u1 is the packet of calls [0, PC), u2 that of
// calls [PC, 2PC); a unit is the sine packet then the cached cosine packet
// (2W outputs), a trailing half unit the sine packet alone.
// MODE 0 standard, 1 z * sd + mean, 3 exp(z * sd + mean).
template <int W, int MODE>
struct NormalF32L {
    using LY = Layout<W>;
    static constexpr int CALLS = 2 * LY::PC, LANES = LY::LANES;
    float* out;
    float mean, sd;
    __device__ void rebase(int64_t off) { out += off; }
    __device__ static int64_t units(int64_t n) { return (n + 2 * W - 1) / (2 * W); }
    __device__ void emit(int64_t per, int64_t u, int l, const uint64_t* w) const {
        const bool full = 2 * W * u + 2 * W <= per;
#pragma unroll
        for (int c = 0; c < LY::PC; ++c) {
            float sn[2], cs[2];
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                box_muller_f32(h ? hi32(w[c]) : lo32(w[c]), h ? hi32(w[LY::PC + c]) : lo32(w[LY::PC + c]), sn[h],
                               cs[h]);
                if (MODE == 1) {
                    sn[h] = sn[h] * sd + mean;
                    cs[h] = cs[h] * sd + mean;
                } else if (MODE == 3) {
                    sn[h] = pexp(sn[h] * sd + mean);
                    cs[h] = pexp(cs[h] * sd + mean);
                }
            }
            float* o = out + 2 * W * u + LY::FPC * c + 2 * l;
            *reinterpret_cast<float2*>(o) = make_float2(sn[0], sn[1]);
            if (full) *reinterpret_cast<float2*>(o + W) = make_float2(cs[0], cs[1]);
        }
    }
};

// StudentTGen (NormalExp.h:285-301) at W = 4 or 16: u1 from calls [0, PC),
// u2 from [PC, 2PC), W outputs per unit, nothing cached.
template <int W>
struct StudentTF32L {
    using LY = Layout<W>;
    static constexpr int CALLS = 2 * LY::PC, LANES = LY::LANES;
    float* out;
    float n, k;  // k = -2/n in float
    __device__ void rebase(int64_t off) { out += off; }
    __device__ static int64_t units(int64_t n_) { return n_ / W; }
    __device__ void emit(int64_t per, int64_t u, int l, const uint64_t* w) const {
#pragma unroll
        for (int c = 0; c < LY::PC; ++c) {
            float v[2];
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const float u1 = 1.0f - ur_f32(h ? hi32(w[c]) : lo32(w[c]));
                const float u2 = ur_f32(h ? hi32(w[LY::PC + c]) : lo32(w[LY::PC + c]));
                const float radius = sqrtf(n * (pexp(plog_pos(u1) * k) - 1.0f));
                float s, cc;
                psincos_pos(ERD_F(2 * kPi) * u2, s, cc);
                v[h] = radius * s;
            }
            *reinterpret_cast<float2*>(out + W * u + LY::FPC * c + 2 * l) = make_float2(v[0], v[1]);
        }
    }
};

// StdNormalGen / NormalGen<double> at W = 4 (Packet2d) or 16 (Packet8d):
// PD = W/2 doubles per packet, LANES per call; u1 from calls [0, PC), u2
// from [PC, 2PC); a unit is 2PD outputs (sine packet, cosine packet).
template <int W, int MODE>
struct NormalF64L {
    static constexpr int LANES = Layout<W>::LANES, PD = W / 2, PC = PD / LANES, CALLS = 2 * PC;
    double* out;
    double mean, sd;
    __device__ void rebase(int64_t off) { out += off; }
    __device__ static int64_t units(int64_t n) { return (n + 2 * PD - 1) / (2 * PD); }
    __device__ void emit(int64_t per, int64_t u, int l, const uint64_t* w) const {
        const bool full = 2 * PD * u + 2 * PD <= per;
#pragma unroll
        for (int c = 0; c < PC; ++c) {
            const double u1 = 1.0 - ur_f64(w[c]);  // in [2^-52, 1], exact
            const double u2 = ur_f64(w[PC + c]);
            const double radius = sqrt(-2.0 * plog_pos_f64(u1));
            double s, cc;
            psincos_f64((2.0 * kPi) * u2, s, cc);
            double sn = radius * s, cs = radius * cc;
            if (MODE == 1) {
                sn = sn * sd + mean;
                cs = cs * sd + mean;
            }
            const int64_t i = 2 * PD * u + LANES * c + l;
            out[i] = sn;
            if (full) out[i + PD] = cs;
        }
    }
};

// ------------------------------------------- raw-bit and VoP float units ---
// One full draw per 8 outputs; float j of the draw comes from 32-bit word j.
// VoP units read per-element parameters laid out like the outputs.

// BernoulliGen<float>::packetOp (Basic.h:693-707): (w & 0x7FFFFFFF) < p with
// p = uint32(p * 2^31) compared as signed int32 lanes (so p = 1 gives 0, as the
// packet path does); BernoulliVGen<float> (Basic.h:722-733) converts
// p * 2^31 with cvttps2dq per element (726-733).
template <int VOP>
struct BernoulliF32 {
    static constexpr int CALLS = 1;
    float* out;
    int32_t p;
    const float* vp;
    __device__ void rebase(int64_t off) {
        out += off;
        if (VOP) vp += off;
    }
    __device__ static int64_t units(int64_t n) { return n / 8; }
    __device__ void emit(int64_t per, int64_t u, int l, const uint64_t* w) const {
        const int64_t i = 8 * u + 2 * l;
        float v[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int32_t r = (int32_t)((h ? hi32(w[0]) : lo32(w[0])) & 0x7FFFFFFFu);
            const int32_t t = VOP ? cvtt_x86(vp[i + h] * 2147483648.0f) : p;
            v[h] = r < t ? 1.0f : 0.0f;
        }
        *reinterpret_cast<float2*>(out + i) = make_float2(v[0], v[1]);
    }
};

// Balanced2Gen (Basic.h:437-470, RandUtils.h:42-45): zero_to_one * slope +
// bias, slope = b - a and bias = a in float; BalancedGen's z*2 - 1 is the
// same arithmetic for (a, b) = (-1, 1). VOP: BalancedVGen (Basic.h:487-493),
// z * (b - a) + a per element.
template <int VOP>
struct BalancedF32 {
    static constexpr int CALLS = 1;
    float* out;
    float slope, bias;
    const float* va;
    const float* vb;
    __device__ void rebase(int64_t off) {
        out += off;
        if (VOP) { va += off; vb += off; }
    }
    __device__ static int64_t units(int64_t n) { return n / 8; }
    __device__ void emit(int64_t per, int64_t u, int l, const uint64_t* w) const {
        const int64_t i = 8 * u + 2 * l;
        float v[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const float z = zero_to_one_f32(h ? hi32(w[0]) : lo32(w[0]));
            v[h] = VOP ? z * (vb[i + h] - va[i + h]) + va[i + h] : z * slope + bias;
        }
        *reinterpret_cast<float2*>(out + i) = make_float2(v[0], v[1]);
    }
};

// Vectorize-over-params units of the single-draw float generators:
//   0 UniformRealVGen   (Basic.h:653-660)      U * (b - a) + a
//   1 ExponentialVGen   (NormalExp.h:418-424)  -(log(1 - U) / a)
//   2 CauchyVGen        (NormalExp.h:875-884)  a + b * tan(pi * (U - 0.5))
//   3 WeibullVGen       (NormalExp.h:675-681)  exp(log(-log(1 - U)) * (1 / a)) * b
//   4 ExtremeValueVGen  (NormalExp.h:752-760)  a - log(-log(U + eps/8)) * b
template <int KIND>
struct VopF32 {
    static constexpr int CALLS = 1;
    float* out;
    const float* va;
    const float* vb;
    __device__ void rebase(int64_t off) {
        out += off;
        va += off;
        if (KIND != 1) vb += off;
    }
    __device__ static int64_t units(int64_t n) { return n / 8; }
    __device__ void emit(int64_t per, int64_t u, int l, const uint64_t* w) const {
        const int64_t i = 8 * u + 2 * l;
        float v[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const uint32_t x = h ? hi32(w[0]) : lo32(w[0]);
            const float a = va[i + h];
            if (KIND == 0) {
                v[h] = ur_f32(x) * (vb[i + h] - a) + a;
            } else if (KIND == 1) {
                v[h] = -(plog_pos(1.0f - ur_f32(x)) / a);
            } else if (KIND == 2) {
                float sn, cs;
                psincos(ERD_F(kPi) * (ur_f32(x) - 0.5f), sn, cs);
                v[h] = a + vb[i + h] * (sn / cs);
            } else if (KIND == 3) {
                v[h] = pexp(plog(-plog(1.0f - ur_f32(x))) * (1.0f / a)) * vb[i + h];
            } else {
                v[h] = a - plog(-plog(nzur_f32(x))) * vb[i + h];
            }
        }
        *reinterpret_cast<float2*>(out + i) = make_float2(v[0], v[1]);
    }
};

// StudentTVGen (NormalExp.h:332-345): StudentTGen with per-element n and -2/n.
struct StudentTVopF32 {
    static constexpr int CALLS = 2;
    float* out;
    const float* vn;
    __device__ void rebase(int64_t off) {
        out += off;
        vn += off;
    }
    __device__ static int64_t units(int64_t n_) { return n_ / 8; }
    __device__ void emit(int64_t per, int64_t u, int l, const uint64_t* w) const {
        const int64_t i = 8 * u + 2 * l;
        float v[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const float n = vn[i + h];
            const float u1 = 1.0f - ur_f32(h ? hi32(w[0]) : lo32(w[0]));
            const float u2 = ur_f32(h ? hi32(w[1]) : lo32(w[1]));
            const float radius = sqrtf(n * (pexp(plog(u1) * (-2.0f / n)) - 1.0f));
            float s, c;
            psincos_pos(ERD_F(2 * kPi) * u2, s, c);
            v[h] = radius * s;
        }
        *reinterpret_cast<float2*>(out + i) = make_float2(v[0], v[1]);
    }
};

// ------------------------------------------------------ half-draw int32 ---
// Packet4i consumers fed by half() (one half draw per 4 outputs, lane j =
// 32-bit word j of the half); emit(i, x) writes outputs i..i+3 from the
// half's tempered words x[0], x[1].

// BernoulliGen<int32_t>::packetOp (Basic.h:693-707) at Packet4i.
struct BernoulliI32 {
    static constexpr int SLOTS = 1;
    int32_t* out;
    int32_t p;
    __device__ void emit(int64_t i, const uint64_t* x) const {
        int4 r;
        r.x = (int32_t)(lo32(x[0]) & 0x7FFFFFFFu) < p ? 1 : 0;
        r.y = (int32_t)(hi32(x[0]) & 0x7FFFFFFFu) < p ? 1 : 0;
        r.z = (int32_t)(lo32(x[1]) & 0x7FFFFFFFu) < p ? 1 : 0;
        r.w = (int32_t)(hi32(x[1]) & 0x7FFFFFFFu) < p ? 1 : 0;
        *reinterpret_cast<int4*>(out + i) = r;
    }
};

// GeometricGen::packetOp (Discrete.h:1372-1380): cvtt(trunc(log(1 - U) *
// rlog_q)) with U from the half draw and rlog_q = 1 / log(1 - p) in float.
struct GeometricI32 {
    static constexpr int SLOTS = 1;
    int32_t* out;
    float rlog_q;
    __device__ int32_t one(uint32_t w) const {
        return cvtt_x86(truncf(plog(1.0f - ur_f32(w)) * rlog_q));
    }
    __device__ void emit(int64_t i, const uint64_t* x) const {
        *reinterpret_cast<int4*>(out + i) =
            make_int4(one(lo32(x[0])), one(hi32(x[0])), one(lo32(x[1])), one(hi32(x[1])));
    }
};

// UniformIntGen<int32_t>::packetOp when pdiff + 1 is a power of two
// (Discrete.h:303-306): (rx & bitmask) + min from one half draw.
struct UniformIntPow2I32 {
    static constexpr int SLOTS = 1;
    int32_t* out;
    uint32_t mask;
    int32_t pmin;
    __device__ int32_t one(uint32_t w) const { return (int32_t)((w & mask) + (uint32_t)pmin); }
    __device__ void emit(int64_t i, const uint64_t* x) const {
        *reinterpret_cast<int4*>(out + i) =
            make_int4(one(lo32(x[0])), one(hi32(x[0])), one(lo32(x[1])), one(hi32(x[1])));
    }
};

// DiscreteGen<int32_t, float>::packetOp (Discrete.h:579-601) at Packet4i.
// K < 16 (SLOTS 1): K - #{c in cdf : U < c} with U from one half draw.
// Otherwise (SLOTS 2): idx = bits & bitmask from one half draw, U from the
// next; idx if U < prob[idx] (float table), else alias[idx].
template <int SLOTS_>
struct DiscreteFI32 {
    static constexpr int SLOTS = SLOTS_;
    int32_t* out;
    const float* cdf;  // or prob
    const int32_t* alias;
    int32_t k;         // cdf size
    uint32_t bitmask;
    __device__ int32_t one(const uint64_t* x, int j) const {
        const uint32_t w0 = (j & 1) ? hi32(x[j >> 1]) : lo32(x[j >> 1]);
        if (SLOTS == 1) {
            const float u = ur_f32(w0);
            int32_t r = k;
            for (int i = 0; i < k; ++i) r -= u < cdf[i] ? 1 : 0;
            return r;
        }
        const uint32_t idx = w0 & bitmask;
        const float u = ur_f32((j & 1) ? hi32(x[2 + (j >> 1)]) : lo32(x[2 + (j >> 1)]));
        return u < cdf[idx] ? (int32_t)idx : alias[idx];
    }
    __device__ void emit(int64_t i, const uint64_t* x) const {
        *reinterpret_cast<int4*>(out + i) = make_int4(one(x, 0), one(x, 1), one(x, 2), one(x, 3));
    }
};

// RandbitsGen<int32_t>::packetOp (Basic.h:394-400) at Packet4i: the half's bits.
struct RandbitsI32 {
    static constexpr int SLOTS = 1;
    int32_t* out;
    __device__ void emit(int64_t i, const uint64_t* x) const {
        *reinterpret_cast<int4*>(out + i) =
            make_int4((int32_t)lo32(x[0]), (int32_t)hi32(x[0]), (int32_t)lo32(x[1]), (int32_t)hi32(x[1]));
    }
};

// GammaGen with integer alpha < 20: -(beta * log(prod of alpha nonzero
// uniforms)), alpha calls per packet (NormalExp.h:567-578).
template <int ALPHA>
struct GammaIntF32 {
    static constexpr int CALLS = ALPHA;
    float* out;
    float beta;
    __device__ void rebase(int64_t off) { out += off; }
    __device__ static int64_t units(int64_t n) { return n / 8; }
    __device__ void emit(int64_t per, int64_t u, int l, const uint64_t* w) const {
        float v[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            float yx = nzur_f32(h ? hi32(w[0]) : lo32(w[0]));
#pragma unroll
            for (int c = 1; c < ALPHA; ++c) yx = yx * nzur_f32(h ? hi32(w[c]) : lo32(w[c]));
            v[h] = -(beta * plog(yx));
        }
        *reinterpret_cast<float2*>(out + 8 * u + 2 * l) = make_float2(v[0], v[1]);
    }
};

// GammaIntF32 with alpha given at run time (for op_fixed_rt): the same
// product of alpha nonzero uniforms in the same order.
struct GammaIntRtF32 {
    float* out;
    float beta;
    int alpha;
    __device__ void rebase(int64_t off) { out += off; }
    __device__ static int64_t units(int64_t n) { return n / 8; }
    template <class F>
    __device__ void emit(int64_t per, int64_t u, int l, F&& fetch) const {
        float v[2];
        const uint64_t w0 = fetch(0);
        float y[2] = {nzur_f32(lo32(w0)), nzur_f32(hi32(w0))};
        for (int c = 1; c < alpha; ++c) {
            const uint64_t w = fetch(c);
            y[0] = y[0] * nzur_f32(lo32(w));
            y[1] = y[1] * nzur_f32(hi32(w));
        }
#pragma unroll
        for (int h = 0; h < 2; ++h) v[h] = -(beta * plog(y[h]));
        *reinterpret_cast<float2*>(out + 8 * u + 2 * l) = make_float2(v[0], v[1]);
    }
};

// BetaGen's gamma-pair regime (NormalExp.h:970-975) when both shapes are
// integers in [1, 20): GammaGen<float>(a, 1) and (b, 1) then take a fixed
// number of full draws (NormalExp.h:561-578: one for shape 1, the exponential;
// `shape` nonzero uniforms multiplied otherwise), so packet k of the fill is
// draws [k(a+b), k(a+b)+a) for p1 and the next b for p2, and packets are
// independent. FisherFGen (ratio != 0) maps x to ratio x / (1 - x).
struct BetaIntF32 {
    float* out;
    int a, b;
    float ratio;
    __device__ void rebase(int64_t off) { out += off; }
    __device__ static int64_t units(int64_t n) { return n / 8; }
    template <class F>
    __device__ void emit(int64_t per, int64_t u, int l, F fetch) const {
        float p[2][2];
        int c = 0;
#pragma unroll
        for (int gi = 0; gi < 2; ++gi) {
            const int shape = gi ? b : a;
            if (shape == 1) {
                const uint64_t w = fetch(c++);
#pragma unroll
                for (int h = 0; h < 2; ++h)
                    p[gi][h] = 1.0f * -(plog_pos(1.0f - ur_f32(h ? hi32(w) : lo32(w))) / 1.0f);
            } else {
                float y[2];
                for (int k = 0; k < shape; ++k) {
                    const uint64_t w = fetch(c++);
#pragma unroll
                    for (int h = 0; h < 2; ++h) {
                        const float v = nzur_f32(h ? hi32(w) : lo32(w));
                        y[h] = k == 0 ? v : y[h] * v;
                    }
                }
#pragma unroll
                for (int h = 0; h < 2; ++h) p[gi][h] = -(1.0f * plog(y[h]));
            }
        }
        float v[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const float x = p[0][h] / (p[0][h] + p[1][h]);
            v[h] = ratio != 0.0f ? (ratio * x) / (1.0f - x) : x;
        }
        *reinterpret_cast<float2*>(out + 8 * u + 2 * l) = make_float2(v[0], v[1]);
    }
};

// ====================================================== rejection policies ===
// W: lanes per packet; SW: words per slot (4 full draw, 2 half draw);
// K: slots per candidate; eval(w, h, v, acc): lane 2p+h of a candidate from
// word p of each of its K slots; store(out, i, v): write output element i.

// W = 8: Packet8f on full draws; W = 4: Packet4f on half draws (the
// GammaGen<float> inside NegativeBinomialGen).
template <int W_>
struct GammaGSW {  // NormalExp.h:536-558
    static constexpr int W = W_, SW = W_ / 2, K = 3;
    using Out = float;
    float beta, px, inv_alpha, am1;
    __device__ void eval(const uint64_t* w, int h, float& v, bool& acc, int) const {
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

template <int W_>
struct GammaTangentW {  // NormalExp.h:580-603
    static constexpr int W = W_, SW = W_ / 2, K = 2;
    using Out = float;
    float beta, sq, am1;
    __device__ void eval(const uint64_t* w, int h, float& v, bool& acc, int) const {
        float ys, yc;
        psincos_pos(ERD_F(kPi) * ur_f32(h ? hi32(w[0]) : lo32(w[0])), ys, yc);
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

using GammaGS = GammaGSW<8>;
using GammaTangent = GammaTangentW<8>;

// BetaGen::packetOp with a < 1 and b < 1 (NormalExp.h:950-968): p1 =
// exp(log U1 * (1/a)), p2 = exp(log U2 * (1/b)), candidate p1 / (p1 + p2),
// accepted if p1 + p2 <= 1. ratio != 0: FisherFGen (NormalExp.h:1017-1022)
// stores ratio * x / (1 - x) with ratio = b/a.
struct BetaJohnk {
    static constexpr int W = 8, SW = 4, K = 2;
    using Out = float;
    float inv_a, inv_b, ratio;
    __device__ void eval(const uint64_t* w, int h, float& v, bool& acc, int) const {
        const float p1 = pexp(plog(ur_f32(h ? hi32(w[0]) : lo32(w[0]))) * inv_a);
        const float p2 = pexp(plog(ur_f32(h ? hi32(w[1]) : lo32(w[1]))) * inv_b);
        const float x = p1 + p2;
        v = p1 / x;
        acc = x <= 1.0f;
    }
    __device__ void store(Out* o, int64_t i, float v) const { o[i] = ratio != 0.0f ? (ratio * v) / (1.0f - v) : v; }
};

// The tangent policies take plgamma_approx of integer arguments from the
// table `lg` (plgamma_int) and skip it where the candidate fails its range
// test: the accept test is false there whatever p2 is.
struct PoissonTangent {  // Discrete.h:875-902 (Packet4f via half())
    static constexpr int W = 4, SW = 2, K = 2;
    using Out = int32_t;
    float psq, pmean, plm, pg1;
    const float* lg;
    __device__ void eval(const uint64_t* w, int h, float& v, bool& acc, int) const {
        float s, c;
        psincos_pos(ERD_F(kPi) * ur_f32(h ? hi32(w[0]) : lo32(w[0])), s, c);
        const float yx = s / c;
        const float fres = truncf(psq * yx + pmean);
        const float p1 = (yx * yx + 1.0f) * ERD_F(0.9);
        const bool c1 = 0.0f <= fres;
        const float lgf = c1 ? plgamma_int(lg, fres + 1.0f) : 0.0f;
        const float p2 = pexp((fres * plm - lgf) - pg1);
        const float u2 = ur_f32(h ? hi32(w[1]) : lo32(w[1]));
        acc = c1 && (u2 <= p1 * p2);
        v = fres;
    }
    __device__ void store(Out* o, int64_t i, float v) const { o[i] = (int32_t)v; }
};

// NegativeBinomialGen's tangent stage (GammaPoisson.h:69-96): the Poisson
// tangent candidate with per-lane constants (sq = sqrt(2 mean), lm = log
// mean, g1 = mean lm - lgamma_approx(mean + 1)), stored with cvttps.
struct NegBinTangent {
    static constexpr int W = 4, SW = 2, K = 2;
    using Out = int32_t;
    const float* mean;  // 4 lanes each
    const float* sq;
    const float* lm;
    const float* g1;
    const float* lg;
    __device__ void eval(const uint64_t* w, int h, float& v, bool& acc, int ln) const {
        float s, c;
        psincos_pos(ERD_F(kPi) * ur_f32(h ? hi32(w[0]) : lo32(w[0])), s, c);
        const float yx = s / c;
        const float fres = truncf(sq[ln] * yx + mean[ln]);
        const float p1 = (yx * yx + 1.0f) * ERD_F(0.9);
        const bool c1 = 0.0f <= fres;
        const float lgf = c1 ? plgamma_int(lg, fres + 1.0f) : 0.0f;
        const float p2 = pexp((fres * lm[ln] - lgf) - g1[ln]);
        const float u2 = ur_f32(h ? hi32(w[1]) : lo32(w[1]));
        acc = c1 && (u2 <= p1 * p2);
        v = fres;
    }
    __device__ void store(Out* o, int64_t i, float v) const { o[i] = cvtt_x86(v); }
};

struct BinomialTangent {  // Discrete.h:1011-1051
    static constexpr int W = 4, SW = 2, K = 2;
    using Out = int32_t;
    float ptrials, psqrt_v, pmean, plp, plq, pg1;
    int32_t trials;
    int flip;  // p != small_p: result is trials - k
    const float* lg;
    __device__ void eval(const uint64_t* w, int h, float& v, bool& acc, int) const {
        float s, c;
        psincos_pos(ERD_F(kPi) * ur_f32(h ? hi32(w[0]) : lo32(w[0])), s, c);
        const float ys = s / c;
        const float fres = truncf(psqrt_v * ys + pmean);
        const float p1 = (ERD_F(1.2) * psqrt_v) * (1.0f + ys * ys);
        const bool c1 = (0.0f <= fres) && (fres <= ptrials);
        float lg1 = 0.0f, lg2 = 0.0f;
        if (c1) {
            lg1 = plgamma_int(lg, fres + 1.0f);
            lg2 = plgamma_int(lg, (ptrials + 1.0f) - fres);
        }
        const float p2 = pexp((((pg1 - lg1) - lg2) + fres * plp) + (ptrials - fres) * plq);
        const float u2 = ur_f32(h ? hi32(w[1]) : lo32(w[1]));
        acc = c1 && (u2 <= p1 * p2);
        v = fres;
    }
    __device__ void store(Out* o, int64_t i, float v) const {
        const int32_t k = (int32_t)v;
        o[i] = flip ? trials - k : k;
    }
};


}  // namespace erd
