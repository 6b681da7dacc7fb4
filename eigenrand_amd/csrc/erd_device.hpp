// Device building blocks for the EigenRand sampling path on gfx950.
//
// Engine model: one Vmt19937_64 (EigenRand/PacketRandomEngine.h:585-599, AVX2
// layout) = 4 independent MT19937-64 lanes. One workgroup of one wave64 owns
// one engine for the whole fill: its 4x312 u64 state lives in LDS, the twist is
// done lane-parallel in three dependency phases, tempering happens in
// registers as words are consumed.
//
// Float math: every function reproduces the op sequence of the reference's
// unfused packet code (AVX2 without FMA), so the library is compiled with
// -ffp-contract=off and correctly rounded fp32 sqrt/div.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#pragma clang fp contract(off)

namespace erd {

constexpr int kN = 312;                 // MT19937-64 state words per lane
constexpr int kM = 156;
constexpr int kLanes = 4;               // u64 lanes per Vmt19937_64 call (AVX2)
constexpr int kWords = kN * kLanes;     // 1248 u64 per engine
constexpr uint64_t kA = 0xb5026f5aa96619e9ull;
constexpr uint64_t kUM = 0xFFFFFFFF80000000ull;
constexpr uint64_t kLM = 0x000000007FFFFFFFull;
constexpr uint64_t kF = 6364136223846793005ull;

constexpr double kPi = 3.1415926535897932;   // EigenRand/Dists/Basic.h:23
constexpr double kE = 2.7182818284590452;    // EigenRand/Dists/Basic.h:24

// Per-engine record in HBM. The layout is part of the C ABI (snapshot/restore).
struct EngineState {
    uint64_t mt[kWords];     // word i of lane l at mt[i*4 + l] (packet-major)
    int32_t idx;             // next word 0..312; 312 = twist before next draw
    int32_t half_valid;      // MersenneTwister::half() cache flag
    uint64_t half_cache[2];  // upper 128 bits of the draw that filled the cache
    uint64_t reserved;
};
static_assert(sizeof(EngineState) == 10016, "engine record layout");

// Cached state of a reused generator object on one engine (Basic.h:268-342):
// StdNormalGen's pending cos packet (count 8) or a rejection sampler's
// CompressMask leftovers (count < packet width). count = 0: nothing cached.
struct GenState {
    int32_t count;
    int32_t reserved;
    float v[8];
};
static_assert(sizeof(GenState) == 40, "generator state layout");

// A per-element parameter array, or one value for every element: the
// reference's mixed overloads pass a scalar as a constant array of the same
// shape (Discrete.h:1852-1897, NormalExp.h:1312-1335), so element i reads v.
template <class T>
struct Param {
    const T* p;  // NULL: every element is v
    T v;
    __device__ T operator[](int64_t i) const { return p ? p[i] : v; }
    __device__ Param operator+(int64_t off) const { return {p ? p + off : nullptr, v}; }
};

// Error bits a kernel can raise (OR-ed into a device status word).
constexpr uint32_t kStatusIterCap = 1u;
constexpr uint32_t kStatusBadSegment = 2u;

// Error reporting and iteration limits of one launch. The reference's
// rejection and lockstep loops run until they succeed (RandUtils.h:342-346);
// here a rejection op gives up after reject_cap candidates without an output
// and a lockstep packet after lockstep_cap iterations, raising kStatusIterCap.
struct Guard {
    uint32_t* status;
    int32_t reject_cap;
    int32_t lockstep_cap;
};
constexpr int32_t kDefaultRejectCap = 1 << 22;
constexpr int32_t kDefaultLockstepCap = 1 << 20;

// Engine (or output block) of workgroup b in a launch of n workgroups, one
// per engine. Workgroups are dealt to the eight XCDs round-robin (b on XCD
// b mod 8); when n is a multiple of 8, XCD x takes the contiguous engines
// [x*n/8, (x+1)*n/8), so each XCD writes one contiguous region. Measured for
// 4096 blocks of 256 KiB: plain 16-byte stores 71% -> 74% of 8 TB/s, the
// fast normal fill 58.5% -> 59.8%.
__device__ __forceinline__ int64_t xcd_block(int64_t b, int64_t n) {
    return (n & 7) ? b : (b & 7) * (n >> 3) + (b >> 3);
}

__device__ __forceinline__ float as_f(uint32_t u) { return __uint_as_float(u); }
__device__ __forceinline__ uint32_t as_u(float f) { return __float_as_uint(f); }

// ----------------------------------------------------------- MT19937-64 ---
// a ^ (b & c) in one v_bitop3_b32 (truth-table index 4a + 2b + c).
__device__ __forceinline__ uint32_t xor_and(uint32_t a, uint32_t b, uint32_t c) {
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x78);
}
// Bits 31..0 of ({hi, lo} >> n), n < 32: one v_alignbit_b32.
__device__ __forceinline__ uint32_t funnel_lo(uint32_t hi, uint32_t lo, uint32_t n) {
    return __builtin_amdgcn_alignbit(hi, lo, n);
}

__device__ __forceinline__ uint64_t temper(uint64_t y) {
    // PacketRandomEngine.h:303-308, on 32-bit halves:
    //   y ^= (y >> 29) & 0x5555555555555555
    //   y ^= (y << 17) & 0x71d67fffeda60000
    //   y ^= (y << 37) & 0xfff7eee000000000
    //   y ^= y >> 43
    uint32_t lo = (uint32_t)y, hi = (uint32_t)(y >> 32);
    const uint32_t a_lo = funnel_lo(hi, lo, 29), a_hi = hi >> 29;
    lo = xor_and(lo, a_lo, 0x55555555u);
    hi = xor_and(hi, a_hi, 0x55555555u);
    const uint32_t b_lo = lo << 17, b_hi = funnel_lo(hi, lo, 15);
    lo = xor_and(lo, b_lo, 0xeda60000u);
    hi = xor_and(hi, b_hi, 0x71d67fffu);
    hi = xor_and(hi, lo << 5, 0xfff7eee0u);
    lo ^= hi >> 11;
    return ((uint64_t)hi << 32) | lo;
}

// Inverse of temper: a word w placed in the state block as untemper(w) is
// read back as w by every consumer (they temper what they read).
__device__ __forceinline__ uint64_t untemper(uint64_t y) {
    y ^= y >> 43;                                  // its own inverse (43 >= 32)
    y ^= (y << 37) & 0xfff7eee000000000ull;        // likewise (37 >= 32)
    uint64_t x = y;
#pragma unroll
    for (int i = 0; i < 3; ++i) x = y ^ ((x << 17) & 0x71d67fffeda60000ull);
    y = x;
#pragma unroll
    for (int i = 0; i < 2; ++i) x = y ^ ((x >> 29) & 0x5555555555555555ull);
    return x;
}

__device__ __forceinline__ uint64_t twist_word(uint64_t cur, uint64_t next, uint64_t far) {
    // y = (cur & UM) | (next & LM); far ^ (y >> 1) ^ (y odd ? A : 0), with
    // UM = upper 33 bits, on 32-bit halves. y's parity is next's (bit 0 is in
    // LM); (y >> 1) ^ (odd & A) is one bitop3 per half.
    const uint32_t y_hi = (uint32_t)(cur >> 32);
    const uint32_t y_lo = __builtin_amdgcn_bitop3_b32((uint32_t)cur, (uint32_t)next, 0x7fffffffu, 0xd8);  // c ? b : a
    const uint32_t sh_lo = funnel_lo(y_hi, y_lo, 1), sh_hi = y_hi >> 1;
    const uint32_t odd = 0u - ((uint32_t)next & 1u);
    const uint32_t r_lo = (uint32_t)far ^ xor_and(sh_lo, odd, (uint32_t)kA);
    const uint32_t r_hi = (uint32_t)(far >> 32) ^ xor_and(sh_hi, odd, (uint32_t)(kA >> 32));
    return ((uint64_t)r_hi << 32) | r_lo;
}

// In-place twist of the 4 lanes held in LDS (packet-major, so consecutive
// threads touch consecutive u64 words: conflict-free ds_read_b64).
// Phase 1 (i < 156) reads only old words; phase 2 (156 <= i < 311) reads the
// phase-1 result at i-156; phase 3 (i = 311) reads new words 0 and 155.
// Within a phase a word i is read (as i+1) before it is rewritten because the
// store depends on the load and later loop trips only touch higher i.
// Twist of words base + lane + 64k, k in [K0, K0 + NT), of a phase with
// `total` words; `far` is the offset of each word's x[i+m] partner. All loads
// of the batch are issued before any of its stores, and the stores before any
// later load: the memory clobbers pin that order, which the other lanes rely
// on (a word is read as a neighbour's `next` before its own lane rewrites it).
template <int K0, int NT>
__device__ __forceinline__ void twist_batch(uint64_t* mt, int base, int total, int far, int lane) {
    uint64_t c[NT], n[NT], f[NT];
#pragma unroll
    for (int i = 0; i < NT; ++i) {
        const int r = lane + 64 * (K0 + i);
        if (64 * (K0 + i) + 63 < total || r < total) {  // full trips need no test
            c[i] = mt[base + r];
            n[i] = mt[base + r + kLanes];
            f[i] = mt[base + r + far];
        }
    }
    asm volatile("" ::: "memory");
#pragma unroll
    for (int i = 0; i < NT; ++i) {
        const int r = lane + 64 * (K0 + i);
        if (64 * (K0 + i) + 63 < total || r < total) mt[base + r] = twist_word(c[i], n[i], f[i]);
    }
    asm volatile("" ::: "memory");
}

// The two large phases run in batches of five 64-word trips, so each batch's
// LDS loads are in flight together instead of one trip at a time.
__device__ __forceinline__ void twist_lds(uint64_t* mt, int lane) {
    constexpr int P1 = kM * kLanes;              // 624 words, trips 0..9
    constexpr int P2 = (kN - 1) * kLanes - P1;   // 620 words, trips 0..9
    twist_batch<0, 5>(mt, 0, P1, P1, lane);
    twist_batch<5, 5>(mt, 0, P1, P1, lane);
    __syncthreads();
    twist_batch<0, 5>(mt, P1, P2, -P1, lane);
    twist_batch<5, 5>(mt, P1, P2, -P1, lane);
    __syncthreads();
    if (lane < kLanes) {
        int t = (kN - 1) * kLanes + lane;
        mt[t] = twist_word(mt[t], mt[lane], mt[t - kM * kLanes]);
    }
    __syncthreads();
}

// twist_lds for a workgroup of NT threads (NT/64 waves), all of which share
// the work: in each of the two large phases every thread loads its words
// (and their next and far partners), the workgroup waits, then every thread
// stores its new words, so each old word is read before any is rewritten.
template <int NT>
__device__ __forceinline__ void twist_lds_wg(uint64_t* mt, int tid) {
    constexpr int P1 = kM * kLanes;              // 624 words
    constexpr int P2 = (kN - 1) * kLanes - P1;   // 620 words
    constexpr int TR = (P1 + NT - 1) / NT;       // words per thread and phase
    uint64_t c[TR], n[TR], f[TR];
#pragma unroll
    for (int i = 0; i < TR; ++i) {
        const int r = tid + NT * i;
        if (r < P1) {
            c[i] = mt[r];
            n[i] = mt[r + kLanes];
            f[i] = mt[r + P1];
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < TR; ++i) {
        const int r = tid + NT * i;
        if (r < P1) mt[r] = twist_word(c[i], n[i], f[i]);
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < TR; ++i) {
        const int r = tid + NT * i;
        if (r < P2) {
            c[i] = mt[P1 + r];
            n[i] = mt[P1 + r + kLanes];
            f[i] = mt[r];  // word r of phase 1, already new
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < TR; ++i) {
        const int r = tid + NT * i;
        if (r < P2) mt[P1 + r] = twist_word(c[i], n[i], f[i]);
    }
    __syncthreads();
    if (tid < kLanes) {
        int t = (kN - 1) * kLanes + tid;
        mt[t] = twist_word(mt[t], mt[tid], mt[t - kM * kLanes]);
    }
    __syncthreads();
}

// Coalesced 16-byte copies of the 9984-byte word block between HBM and LDS.
__device__ __forceinline__ void load_state(uint64_t* lds, const EngineState* st, int lane) {
    const uint4* src = reinterpret_cast<const uint4*>(st->mt);
    uint4* dst = reinterpret_cast<uint4*>(lds);
    for (int i = lane; i < kWords / 2; i += 64) dst[i] = src[i];
    __syncthreads();
}
__device__ __forceinline__ void store_state(EngineState* st, const uint64_t* lds, int lane) {
    const uint4* src = reinterpret_cast<const uint4*>(lds);
    uint4* dst = reinterpret_cast<uint4*>(st->mt);
    for (int i = lane; i < kWords / 2; i += 64) dst[i] = src[i];
}

// ------------------------------------------------------ bits -> uniforms ---
// bit_to_ur_float: MorePacketMath.h:205-215.
// Correctly rounded sqrt for x = +-0 or x >= 2^-96 (normal, no rescaling
// needed): v_sqrt_f32 is within 1 ulp; the neighbour whose residual shows it
// is the rounded root replaces it (the same correction the compiler applies
// for -fhip-fp32-correctly-rounded-divide-sqrt, without its denormal scaling
// and zero/infinity checks, which these inputs never need).
__device__ __forceinline__ float sqrt_cr_normal(float x) {
    const float s = __builtin_amdgcn_sqrtf(x);
    const float dn = __int_as_float(__float_as_int(s) - 1);
    const float up = __int_as_float(__float_as_int(s) + 1);
    const float r_dn = __builtin_fmaf(-dn, s, x);
    const float r_up = __builtin_fmaf(-up, s, x);
    float r = (r_dn <= 0.0f) ? dn : s;
    r = (r_up > 0.0f) ? up : r;
    return r;
}

// 1 - bit_to_ur_float(w): (f - 1) is exact for f in [1, 2), and so is
// 1 - (f - 1) = 2 - f.
__device__ __forceinline__ float one_minus_ur_f32(uint32_t w) {
    return 2.0f - as_f((w & 0x7FFFFFu) | 0x3F800000u);
}

__device__ __forceinline__ float ur_f32(uint32_t w) {
    return as_f((w & 0x7FFFFFu) | 0x3F800000u) - 1.0f;
}
// nonzero_uniform_real: RandUtils.h:47-51 (epsilon/8).
__device__ __forceinline__ float nzur_f32(uint32_t w) {
    return ur_f32(w) + 1.1920928955078125e-07f / 8;
}
// RandUtils::zero_to_one at Packet8f (arch/AVX/RandUtils.h:200-204):
// cvtepi32_ps(w & 0x7FFFFFFF) / float(0x7FFFFFFF); the divisor rounds to 2^31,
// so the quotient is the converted integer scaled exactly.
__device__ __forceinline__ float zero_to_one_f32(uint32_t w) {
    return (float)(int32_t)(w & 0x7FFFFFFFu) / 2147483648.0f;
}
// x86 cvttps2dq: truncation, with NaN and out-of-range inputs giving INT32_MIN.
__device__ __forceinline__ int32_t cvtt_x86(float f) {
    if (!(f < 2147483648.0f) || !(f >= -2147483648.0f)) return (int32_t)0x80000000u;
    return (int32_t)f;
}
// bit_to_ur_double: MorePacketMath.h:217-227.
__device__ __forceinline__ double ur_f64(uint64_t w) {
    return __longlong_as_double((long long)((w & 0xFFFFFFFFFFFFFull) | (1023ull << 52))) - 1.0;
}

// ------------------------------------------------------------ float math ---
#define ERD_F(x) ((float)(x))

// ln(x): cephes logf with ln2 split in q1/q2, as the reference writes it
// (EigenRand/arch/NEON/MorePacketMath.h:440-521), unfused.
__device__ __forceinline__ float plog(float x0) {
    const bool invalid = !(0.0f <= x0);
    const bool iszero = (x0 == 0.0f);
    const float min_norm = as_f(0x00800000u);
    float x = (x0 < min_norm) ? min_norm : x0;
    int32_t emm0 = (int32_t)(as_u(x) >> 23) - 0x7f;
    x = as_f((as_u(x) & 0x807FFFFFu) | 0x3F000000u);
    float e = (float)emm0 + 1.0f;
    const bool mask = x < 0.707106781186547524f;
    float tmp = mask ? x : 0.0f;
    x = x - 1.0f;
    e = e - (mask ? 1.0f : 0.0f);
    x = x + tmp;
    const float x2 = x * x;
    const float x3 = x2 * x;
    float y = 7.0376836292E-2f * x + -1.1514610310E-1f;
    float y1 = -1.2420140846E-1f * x + 1.4249322787E-1f;
    float y2 = 2.0000714765E-1f * x + -2.4999993993E-1f;
    y = y * x + 1.1676998740E-1f;
    y1 = y1 * x + -1.6668057665E-1f;
    y2 = y2 * x + 3.3333331174E-1f;
    y = y * x3 + y1;
    y = y * x3 + y2;
    y = y * x3;
    y1 = e * -2.12194440e-4f;
    tmp = x2 * 0.5f;
    y = y + y1;
    x = x - tmp;
    y2 = e * 0.693359375f;
    x = x + y;
    x = x + y2;
    if (iszero) return as_f(0xff800000u);
    if (invalid) return as_f(0xFFFFFFFFu);
    return x;
}

// plog restricted to inputs in [2^-126, +inf) finite: the special-case masks
// of plog are constant there, so the result is identical and cheaper. For such
// x, frexp gives the reference's mantissa with the exponent set to -1 and the
// exponent emm0 + 1, so e is formed from an integer in one conversion.
__device__ __forceinline__ float plog_pos(float x) {
    const int32_t k = __builtin_amdgcn_frexp_expf(x);
    x = __builtin_amdgcn_frexp_mantf(x);
    const bool mask = x < 0.707106781186547524f;
    const float e = (float)(mask ? k - 1 : k);
    float tmp = mask ? x : 0.0f;
    x = x - 1.0f;
    x = x + tmp;
    const float x2 = x * x;
    const float x3 = x2 * x;
    float y = 7.0376836292E-2f * x + -1.1514610310E-1f;
    float y1 = -1.2420140846E-1f * x + 1.4249322787E-1f;
    float y2 = 2.0000714765E-1f * x + -2.4999993993E-1f;
    y = y * x + 1.1676998740E-1f;
    y1 = y1 * x + -1.6668057665E-1f;
    y2 = y2 * x + 3.3333331174E-1f;
    y = y * x3 + y1;
    y = y * x3 + y2;
    y = y * x3;
    y1 = e * -2.12194440e-4f;
    tmp = x2 * 0.5f;
    y = y + y1;
    x = x - tmp;
    y2 = e * 0.693359375f;
    x = x + y;
    return x + y2;
}

// exp(x): cephes expf (ln2 split C1/C2, degree-5 polynomial, scale by 2^m),
// unfused; same op sequence as the oracle's pexp.
__device__ __forceinline__ float pexp(float x0) {
    float x = x0 < 88.723f ? x0 : 88.723f;
    x = x > -88.723f ? x : -88.723f;
    const float m = floorf(x * 1.44269504088896341f + 0.5f);
    float r = m * -0.693359375f + x;
    r = m * 2.12194440e-4f + r;
    const float r2 = r * r;
    const float r3 = r2 * r;
    float y = 1.9875691500E-4f * r + 1.3981999507E-3f;
    float y1 = 4.1665795894E-2f * r + 1.6666665459E-1f;
    const float y2 = r + 1.0f;
    y = y * r + 8.3334519073E-3f;
    y1 = y1 * r + 5.0000001201E-1f;
    y = y * r3 + y1;
    y = y * r2 + y2;
    const float out = ldexpf(y, (int)m);
    if (!(x0 == x0)) return x0;
    return out > x0 ? out : x0;
}

// sincos: the reference's cephes scheme, MorePacketMath.h:297-393 (float).
__device__ __forceinline__ void psincos(float xin, float& s, float& c) {
    uint32_t sign_sin = as_u(xin) & 0x80000000u;
    float x = as_f(as_u(xin) & 0x7FFFFFFFu);
    float y = x * ERD_F(1.27323954473516);
    int32_t j = (int32_t)y;
    j = (j + 1) & ~1;
    y = (float)j;
    const uint32_t swap_sin = (uint32_t)(j & 4) << 29;
    const bool poly_mask = (j & 2) == 0;
    x = x + y * ERD_F(-0.78515625);
    x = x + y * ERD_F(-2.4187564849853515625e-4);
    x = x + y * ERD_F(-3.77489497744594108e-8);
    const uint32_t sign_cos = (uint32_t)(4 & ~(j - 2)) << 29;
    sign_sin ^= swap_sin;
    const float z = x * x;
    float yc = ERD_F(2.443315711809948E-005) * z;
    yc = yc + ERD_F(-1.388731625493765E-003);
    yc = yc * z;
    yc = yc + ERD_F(4.166664568298827E-002);
    yc = yc * z;
    yc = yc * z;
    yc = yc - z * ERD_F(0.5);
    yc = yc + 1.0f;
    float ys = ERD_F(-1.9515295891E-4) * z;
    ys = ys + ERD_F(8.3321608736E-3);
    ys = ys * z;
    ys = ys + ERD_F(-1.6666654611E-1);
    ys = ys * z;
    ys = ys * x;
    ys = ys + x;
    // The reference blends with and/andnot then adds; for finite polynomial
    // values that equals a select with -0 mapped to +0 (adding a masked +0).
    const float rs = (poly_mask ? ys : yc) + 0.0f;
    const float rc = (poly_mask ? yc : ys) + 0.0f;
    s = as_f(as_u(rs) ^ sign_sin);
    c = as_f(as_u(rc) ^ sign_cos);
}

// psincos for x >= +0 (the callers pass pi*U or 2*pi*U with U in [0, 1)).
// The input's sign bit is 0, and the reduced argument never becomes -0 (each
// reduction step adds a nonpositive term to a value that is not -0, and an
// exact cancellation rounds to +0), so neither polynomial can yield -0 and the
// reference's masked "+0" addition is an identity. Both are left out; the
// results equal psincos's bit for bit on this domain.
__device__ __forceinline__ void psincos_pos(float x, float& s, float& c) {
    float y = x * ERD_F(1.27323954473516);
    int32_t j = (int32_t)y;
    j = (j + 1) & ~1;
    y = (float)j;
    const uint32_t sign_sin = (uint32_t)(j & 4) << 29;
    const bool poly_mask = (j & 2) == 0;
    x = x + y * ERD_F(-0.78515625);
    x = x + y * ERD_F(-2.4187564849853515625e-4);
    x = x + y * ERD_F(-3.77489497744594108e-8);
    const uint32_t sign_cos = (uint32_t)(4 & ~(j - 2)) << 29;
    const float z = x * x;
    float yc = ERD_F(2.443315711809948E-005) * z;
    yc = yc + ERD_F(-1.388731625493765E-003);
    yc = yc * z;
    yc = yc + ERD_F(4.166664568298827E-002);
    yc = yc * z;
    yc = yc * z;
    yc = yc - z * ERD_F(0.5);
    yc = yc + 1.0f;
    float ys = ERD_F(-1.9515295891E-4) * z;
    ys = ys + ERD_F(8.3321608736E-3);
    ys = ys * z;
    ys = ys + ERD_F(-1.6666654611E-1);
    ys = ys * z;
    ys = ys * x;
    ys = ys + x;
    s = as_f(as_u(poly_mask ? ys : yc) ^ sign_sin);
    c = as_f(as_u(poly_mask ? yc : ys) ^ sign_cos);
}

// ------------------------------------------------------------ double math ---
__device__ __forceinline__ double as_d(uint64_t u) { return __longlong_as_double((long long)u); }
__device__ __forceinline__ uint64_t as_u64(double d) { return (uint64_t)__double_as_longlong(d); }

// ln(x) for finite x >= 2^-1022: the reference's Packet4d plog
// (arch/AVX/MorePacketMath.h:781-858), i.e. the single-precision cephes
// polynomial evaluated in double with ln2 split in q1/q2, unfused. The
// exponent is converted exactly (uint64_to_double, 696-699) and the
// zero/negative/NaN masks are constant on this domain.
__device__ __forceinline__ double plog_pos_f64(double x) {
    const uint64_t b = as_u64(x);
    double e = (double)(b >> 52) - 1022.0;
    x = as_d((b & ~0x7ff0000000000000ull) | 0x3FE0000000000000ull);
    const bool mask = x < 0.707106781186547524;
    double tmp = mask ? x : 0.0;
    x = x - 1.0;
    e = e - (mask ? 1.0 : 0.0);
    x = x + tmp;
    const double x2 = x * x;
    const double x3 = x2 * x;
    double y = 7.0376836292E-2 * x + -1.1514610310E-1;
    double y1 = -1.2420140846E-1 * x + 1.4249322787E-1;
    double y2 = 2.0000714765E-1 * x + -2.4999993993E-1;
    y = y * x + 1.1676998740E-1;
    y1 = y1 * x + -1.6668057665E-1;
    y2 = y2 * x + 3.3333331174E-1;
    y = y * x3 + y1;
    y = y * x3 + y2;
    y = y * x3;
    y1 = e * -2.12194440e-4;
    tmp = x2 * 0.5;
    y = y + y1;
    x = x - tmp;
    y2 = e * 0.693359375;
    x = x + y;
    return x + y2;
}

// sincos for doubles: the reference's cephes scheme at double packets
// (MorePacketMath.h:396-491), single-precision coefficients evaluated in
// double. The quadrant j = floor(|x| 4/pi) goes through the reference's
// double_to_int64 (arch/AVX/MorePacketMath.h:731-737: floor, then the
// 1.5*2^52 bias), which is exact for |x| 4/pi < 2^51.
__device__ __forceinline__ void psincos_f64(double xin, double& s, double& c) {
    uint64_t sign_sin = as_u64(xin) & 0x8000000000000000ull;
    double x = as_d(as_u64(xin) & 0x7FFFFFFFFFFFFFFFull);
    double y = x * 1.27323954473516;
    int64_t j = (int64_t)floor(y);
    j = (j + 1) & ~(int64_t)1;
    y = (double)j;
    const uint64_t swap_sin = (uint64_t)(j & 4) << 61;
    const bool poly_mask = (j & 2) == 0;
    x = x + y * -0.78515625;
    x = x + y * -2.4187564849853515625e-4;
    x = x + y * -3.77489497744594108e-8;
    const uint64_t sign_cos = (uint64_t)(4 & ~(j - 2)) << 61;
    sign_sin ^= swap_sin;
    const double z = x * x;
    double yc = 2.443315711809948E-005 * z;
    yc = yc + -1.388731625493765E-003;
    yc = yc * z;
    yc = yc + 4.166664568298827E-002;
    yc = yc * z;
    yc = yc * z;
    yc = yc - z * 0.5;
    yc = yc + 1.0;
    double ys = -1.9515295891E-4 * z;
    ys = ys + 8.3321608736E-3;
    ys = ys * z;
    ys = ys + -1.6666654611E-1;
    ys = ys * z;
    ys = ys * x;
    ys = ys + x;
    // masked select plus the masked +0, as in the float form
    const double rs = (poly_mask ? ys : yc) + 0.0;
    const double rc = (poly_mask ? yc : ys) + 0.0;
    s = as_d(as_u64(rs) ^ sign_sin);
    c = as_d(as_u64(rc) ^ sign_cos);
}

// lgamma approximation: MorePacketMath.h:158-170.
__device__ __forceinline__ float plgamma_approx(float x) {
    const float x_3 = x + 3.0f;
    float ret = (x_3 + -0.5f) * plog(x_3);
    ret = ret - x_3;
    ret = ret + ERD_F(0.9189385332046727);
    ret = ret + ERD_F(1 / 12.) / x_3;
    ret = ret - plog(((x_3 - 1.0f) * (x_3 - 2.0f)) * x);
    return ret;
}

// plgamma_approx at integer-valued arguments. The library keeps, per device,
// a table of plgamma_approx(k) for k = 1..kLgTab computed by plgamma_approx
// itself (so the values are identical); other arguments are computed. The
// tangent samplers evaluate it only at fres + 1 and trials + 1 - fres, which
// are integer-valued (fres = trunc(...)) or not finite.
constexpr int kLgTab = 1 << 20;
__device__ __forceinline__ float plgamma_int(const float* lg, float x) {
    if (x >= 1.0f && x <= (float)kLgTab) return lg[(int)x - 1];
    return plgamma_approx(x);
}

__device__ __forceinline__ uint32_t lo32(uint64_t w) { return (uint32_t)w; }
__device__ __forceinline__ uint32_t hi32(uint64_t w) { return (uint32_t)(w >> 32); }

}  // namespace erd
