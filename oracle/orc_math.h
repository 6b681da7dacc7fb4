// Oracle (test infrastructure only): scalar restatement of the float packet
// math that the reference's generators call. Every function here works on one
// lane; a packet op in the reference is the same scalar op applied lane by
// lane, so lane emulation reproduces the packet result bit for bit provided
// the build performs no FP contraction (compile with -ffp-contract=off) and
// uses SSE float arithmetic (x86-64 default), i.e. IEEE binary32 with
// round-to-nearest and correctly rounded sqrt/div.
//
// Only tests/, the smoke check and the CPU baseline leg of the benchmark may
// use this code; it is never on the product path.
#pragma once
#include <cstdint>
#include <cstring>
#include <cmath>

namespace orc {

static inline float as_f(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }
static inline uint32_t as_u(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }
static inline double as_d(uint64_t u) { double f; std::memcpy(&f, &u, 8); return f; }

// pset1<Packet>(double literal): the double is rounded once to float.
#define F(x) ((float)(x))

// The reference's numeric constants (EigenRand/Dists/Basic.h:23-24), doubles.
static constexpr double kPi = 3.1415926535897932;
static constexpr double kE = 2.7182818284590452;

// bit_to_ur_float: EigenRand/MorePacketMath.h:205-215 (BitScalar<float>::to_ur
// 241-250 is the same map). 23 mantissa bits -> [1,2) -> minus one.
static inline float bit_to_ur_float(uint32_t w) {
    return as_f((w & 0x7FFFFFu) | 0x3F800000u) - 1.0f;
}
// bit_to_ur_double: EigenRand/MorePacketMath.h:217-227.
static inline double bit_to_ur_double(uint64_t w) {
    return as_d((w & 0xFFFFFFFFFFFFFull) | (1023ull << 52)) - 1.0;
}
// RandUtils::nonzero_uniform_real adds epsilon/8: EigenRand/RandUtils.h:47-51.
static inline float nzur_float(uint32_t w) {
    return bit_to_ur_float(w) + 1.1920928955078125e-07f / 8;
}

// Natural log, cephes single-precision scheme with the ln2 split into
// q1/q2, exactly as written in the reference (EigenRand/arch/NEON/
// MorePacketMath.h:440-521). pmadd(a,b,c) is an unfused multiply then add,
// which is what Eigen's generic pmadd does on an AVX2 build without FMA.
static inline float plog(float x0) {
    float x = x0;
    const bool invalid = !(0.0f <= x);            // pbitnot(pcmple(0, x)), NaN too
    const bool iszero = (x == 0.0f);
    const float min_norm = as_f(0x00800000u);
    x = (x < min_norm) ? min_norm : x;            // pmax(x, min_norm); NaN/negative inputs are overridden below
    int32_t emm0 = (int32_t)(as_u(x) >> 23);
    x = as_f((as_u(x) & ~0x7f800000u) | as_u(0.5f));
    emm0 -= 0x7f;
    float e = (float)emm0 + 1.0f;
    const bool mask = x < 0.707106781186547524f;
    float tmp = mask ? x : 0.0f;
    x = x - 1.0f;
    e = e - (mask ? 1.0f : 0.0f);
    x = x + tmp;
    const float x2 = x * x;
    const float x3 = x2 * x;
    float y, y1, y2;
    y  = 7.0376836292E-2f * x + -1.1514610310E-1f;
    y1 = -1.2420140846E-1f * x + +1.4249322787E-1f;
    y2 = +2.0000714765E-1f * x + -2.4999993993E-1f;
    y  = y * x + 1.1676998740E-1f;
    y1 = y1 * x + -1.6668057665E-1f;
    y2 = y2 * x + +3.3333331174E-1f;
    y  = y * x3 + y1;
    y  = y * x3 + y2;
    y  = y * x3;
    y1 = e * -2.12194440e-4f;
    tmp = x2 * 0.5f;
    y = y + y1;
    x = x - tmp;
    y2 = e * 0.693359375f;
    x = x + y;
    x = x + y2;
    if (iszero) return as_f(0xff800000u);
    if (invalid) return as_f(as_u(x) | 0xFFFFFFFFu);
    return x;
}

// exp, cephes single-precision scheme (range reduction by ln2 split into
// C1/C2, degree-5 polynomial, scale by 2^m), written from the published cephes
// expf method in the form Eigen 3.4 evaluates it (unfused pmadd). The
// reference calls Eigen's pexp here; its exact bits are not in the reference
// text, so this function is pinned against libm by an ulp bound instead.
static inline float pexp(float x0) {
    float x = x0;
    x = x < 88.723f ? x : 88.723f;
    x = x > -88.723f ? x : -88.723f;
    if (!(x0 == x0)) x = x0;
    float m = std::floor(x * 1.44269504088896341f + 0.5f);
    float r = m * -0.693359375f + x;
    r = m * 2.12194440e-4f + r;
    const float r2 = r * r;
    const float r3 = r2 * r;
    float y, y1, y2;
    y  = 1.9875691500E-4f * r + 1.3981999507E-3f;
    y1 = 4.1665795894E-2f * r + 1.6666665459E-1f;
    y2 = r + 1.0f;
    y  = y * r + 8.3334519073E-3f;
    y1 = y1 * r + 5.0000001201E-1f;
    y  = y * r3 + y1;
    y  = y * r2 + y2;
    // pldexp(y, m) = y * 2^m; for m >= -128 Eigen's staged scaling keeps every
    // intermediate normal, so it rounds once, exactly as ldexp does.
    int32_t mi = (int32_t)m;
    float out = std::ldexp(y, mi);
    if (!(x0 == x0)) return x0;
    return out > x0 ? out : x0;                   // pmax(pldexp(y, m), x)
}

// sincos, the reference's own cephes scheme: EigenRand/MorePacketMath.h:297-393
// (float packets). pnew_andnot(a, b) = a & ~b (MorePacketMath.h:229-236).
static inline void psincos(float xin, float& s, float& c) {
    float x = xin;
    uint32_t sign_bit_sin = as_u(x) & 0x80000000u;
    x = as_f(as_u(x) & 0x7FFFFFFFu);
    float y = x * F(1.27323954473516);
    int32_t emm2 = (int32_t)y;                    // cvttps: truncation
    emm2 = emm2 + 1;
    emm2 = emm2 & ~1;
    y = (float)emm2;
    int32_t emm4 = emm2;
    uint32_t swap_sign_bit_sin = (uint32_t)(emm2 & 4) << 29;
    const bool poly_mask = (emm2 & 2) == 0;
    x = x + y * F(-0.78515625);
    x = x + y * F(-2.4187564849853515625e-4);
    x = x + y * F(-3.77489497744594108e-8);
    emm4 = emm4 - 2;
    emm4 = 4 & ~emm4;
    uint32_t sign_bit_cos = (uint32_t)emm4 << 29;
    sign_bit_sin ^= swap_sign_bit_sin;
    const float z = x * x;
    float yc = F(2.443315711809948E-005);
    yc = yc * z;
    yc = yc + F(-1.388731625493765E-003);
    yc = yc * z;
    yc = yc + F(4.166664568298827E-002);
    yc = yc * z;
    yc = yc * z;
    const float tmp = z * F(0.5);
    yc = yc - tmp;
    yc = yc + F(1);
    float ys = F(-1.9515295891E-4);
    ys = ys * z;
    ys = ys + F(8.3321608736E-3);
    ys = ys * z;
    ys = ys + F(-1.6666654611E-1);
    ys = ys * z;
    ys = ys * x;
    ys = ys + x;
    // blend by masks then add: ysin2 = mask&ys, ysin1 = yc&~mask
    const float ysin2 = poly_mask ? ys : 0.0f;
    const float ysin1 = poly_mask ? 0.0f : yc;
    ys = ys - ysin2;
    yc = yc - ysin1;
    const float xmm1 = ysin1 + ysin2;
    const float xmm2 = yc + ys;
    s = as_f(as_u(xmm1) ^ sign_bit_sin);
    c = as_f(as_u(xmm2) ^ sign_bit_cos);
}

static inline uint64_t as_u64(double d) { uint64_t u; std::memcpy(&u, &d, 8); return u; }

// The reference's 64-bit lane conversions (arch/AVX/MorePacketMath.h:696-737),
// written with their bias constants: uint64_to_double ORs in the bits of 2^52
// and subtracts 2^52; int64_to_double adds the bits of 1.5*2^52 as an integer
// and subtracts 1.5*2^52; double_to_int64 floors, adds 1.5*2^52 and
// subtracts the bias bits as integers.
static inline double uint64_to_double(uint64_t x) {
    return as_d(x | as_u64(4503599627370496.0)) - 4503599627370496.0;
}
static inline double int64_to_double(int64_t x) {
    return as_d((uint64_t)x + as_u64(6755399441055744.0)) - 6755399441055744.0;
}
static inline int64_t double_to_int64(double x) {
    x = std::floor(x) + 6755399441055744.0;
    return (int64_t)(as_u64(x) - as_u64(6755399441055744.0));
}

// plog<Packet4d>, the reference's double log (arch/AVX/MorePacketMath.h:
// 781-858): single-precision cephes coefficients evaluated in double with ln2
// split in q1/q2, unfused; 0 -> -inf, negative or NaN -> NaN (all bits set).
static inline double plog_d(double x0) {
    double x = x0;
    const bool invalid = !(x >= 0.0);             // _CMP_NGE_UQ
    const bool iszero = (x == 0.0);
    const double min_norm = as_d(0x10000000000000ull);
    x = (x > min_norm) ? x : min_norm;            // _mm256_max_pd(x, min_norm)
    const double emm0 = uint64_to_double(as_u64(x) >> 52);
    double e = emm0 - 1022.0;
    x = as_d((as_u64(x) & ~0x7ff0000000000000ull) | as_u64(0.5));
    const bool mask = x < 0.707106781186547524;
    double tmp = mask ? x : 0.0;
    x = x - 1.0;
    e = e - (mask ? 1.0 : 0.0);
    x = x + tmp;
    const double x2 = x * x;
    const double x3 = x2 * x;
    double y, y1, y2;
    y = 7.0376836292E-2 * x + -1.1514610310E-1;
    y1 = -1.2420140846E-1 * x + +1.4249322787E-1;
    y2 = +2.0000714765E-1 * x + -2.4999993993E-1;
    y = y * x + 1.1676998740E-1;
    y1 = y1 * x + -1.6668057665E-1;
    y2 = y2 * x + +3.3333331174E-1;
    y = y * x3 + y1;
    y = y * x3 + y2;
    y = y * x3;
    y1 = e * -2.12194440e-4;
    tmp = x2 * 0.5;
    y = y + y1;
    x = x - tmp;
    y2 = e * 0.693359375;
    x = x + y;
    x = x + y2;
    if (iszero) return as_d(0xfff0000000000000ull);
    if (invalid) return as_d(as_u64(x) | 0xFFFFFFFFFFFFFFFFull);
    return x;
}

// sincos at double packets: EigenRand/MorePacketMath.h:396-491 (the float
// scheme's steps with 64-bit lanes and the conversions above).
static inline void psincos_d(double xin, double& s, double& c) {
    double x = xin;
    uint64_t sign_bit_sin = as_u64(x) & 0x8000000000000000ull;
    x = as_d(as_u64(x) & 0x7FFFFFFFFFFFFFFFull);
    double y = x * 1.27323954473516;
    int64_t emm2 = double_to_int64(y);
    emm2 = emm2 + 1;
    emm2 = emm2 & ~1ll;
    y = int64_to_double(emm2);
    int64_t emm4 = emm2;
    const uint64_t swap_sign_bit_sin = (uint64_t)(emm2 & 4) << 61;
    const bool poly_mask = (emm2 & 2) == 0;
    const double xmm1 = y * -0.78515625;
    const double xmm2 = y * -2.4187564849853515625e-4;
    const double xmm3 = y * -3.77489497744594108e-8;
    x = x + xmm1;
    x = x + xmm2;
    x = x + xmm3;
    emm4 = emm4 - 2;
    emm4 = 4 & ~emm4;
    const uint64_t sign_bit_cos = (uint64_t)emm4 << 61;
    sign_bit_sin ^= swap_sign_bit_sin;
    const double z = x * x;
    double yc = 2.443315711809948E-005;
    yc = yc * z;
    yc = yc + -1.388731625493765E-003;
    yc = yc * z;
    yc = yc + 4.166664568298827E-002;
    yc = yc * z;
    yc = yc * z;
    const double tmp = z * 0.5;
    yc = yc - tmp;
    yc = yc + 1.0;
    double ys = -1.9515295891E-4;
    ys = ys * z;
    ys = ys + 8.3321608736E-3;
    ys = ys * z;
    ys = ys + -1.6666654611E-1;
    ys = ys * z;
    ys = ys * x;
    ys = ys + x;
    const double ysin2 = poly_mask ? ys : 0.0;
    const double ysin1 = poly_mask ? 0.0 : yc;
    ys = ys - ysin2;
    yc = yc - ysin1;
    const double r1 = ysin1 + ysin2;
    const double r2 = yc + ys;
    s = as_d(as_u64(r1) ^ sign_bit_sin);
    c = as_d(as_u64(r2) ^ sign_bit_cos);
}

// plgamma_approx: EigenRand/MorePacketMath.h:158-170 (float constants).
static inline float plgamma_approx(float x) {
    const float x_3 = x + 3.0f;
    float ret = (x_3 + (float)-0.5) * plog(x_3);
    ret = ret - x_3;
    ret = ret + (float)0.9189385332046727;
    ret = ret + (float)(1 / 12.) / x_3;
    ret = ret - plog(((x_3 - 1.0f) * (x_3 - 2.0f)) * x);
    return ret;
}

// ptruncate: round toward zero (arch/AVX/MorePacketMath.h:651-655).
static inline float ptruncate(float a) { return std::trunc(a); }

// pcast<Packet4f/8f, Packet4i/8i> = cvttps2dq: truncation; NaN and values
// outside the int32 range give 0x80000000 (the "integer indefinite" value).
static inline int32_t cvttps(float f) {
    if (!(f < 2147483648.0f) || !(f >= -2147483648.0f)) return INT32_MIN;
    return (int32_t)f;
}

// RandUtils<Packet8f>::zero_to_one (arch/AVX/RandUtils.h:200-204):
// cvtepi32_ps(w & 0x7FFFFFFF) / pset1<Packet8f>(0x7FFFFFFF).
static inline float zero_to_one_float(uint32_t w) {
    return (float)(int32_t)(w & 0x7FFFFFFFu) / (float)0x7FFFFFFF;
}

}  // namespace orc
