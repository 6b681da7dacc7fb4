// Fills whose consumption per output packet is fixed: raw u64, uniformReal
// (f32/f64), normal (Box-Muller pairs), exponential, Cauchy, Student-t, gamma
// with alpha == 1 or integer alpha < 20, and NormalVGen (per-element mean/sd).
//
// One workgroup = one wave64 = one Vmt19937_64 engine. The engine's 312x4 u64
// state is loaded into LDS once, twisted in place block by block, and written
// back only if a twist happened. A "unit" is the group of engine calls that
// produces one output packet (or a Box-Muller packet pair); each thread takes
// (unit, lane) items, so every MT word is read and tempered exactly once.
// Units that straddle a block boundary keep their first words in a small LDS
// carry area across the twist.
#include "erd_device.hpp"
#include "erd_kernels.hpp"

namespace erd {

// ---------------------------------------------------------------- units ---
// Each functor: CALLS engine calls per unit; units(n) = units for n outputs;
// emit(...) writes the outputs of one (unit, lane) item.

struct RawU64 {  // out[4u + l] = lane l of call u (PacketRandomEngine.h:295-310)
    static constexpr int CALLS = 1;
    uint64_t* out;
    __device__ static int64_t units(int64_t n) { return n / 4; }
    __device__ void emit(int64_t g, int64_t per, int64_t u, int l, const uint64_t* w) const {
        out[g * per + 4 * u + l] = w[0];
    }
};

struct UniformF32 {  // Basic.h:584-590 / 626-634; float j of a call = word j
    static constexpr int CALLS = 1;
    float* out;
    float a, slope;
    int affine;
    __device__ static int64_t units(int64_t n) { return n / 8; }
    __device__ void emit(int64_t g, int64_t per, int64_t u, int l, const uint64_t* w) const {
        float u0 = ur_f32(lo32(w[0])), u1 = ur_f32(hi32(w[0]));
        if (affine) { u0 = u0 * slope + a; u1 = u1 * slope + a; }
        *reinterpret_cast<float2*>(out + g * per + 8 * u + 2 * l) = make_float2(u0, u1);
    }
};

struct UniformF64 {  // arch/AVX/RandUtils.h:212-225: 4 doubles per call
    static constexpr int CALLS = 1;
    double* out;
    double a, slope;
    int affine;
    __device__ static int64_t units(int64_t n) { return n / 4; }
    __device__ void emit(int64_t g, int64_t per, int64_t u, int l, const uint64_t* w) const {
        double v = ur_f64(w[0]);
        if (affine) v = v * slope + a;
        out[g * per + 4 * u + l] = v;
    }
};

// Box-Muller pair of StdNormalGen::packetOp (NormalExp.h:60-82): calls A, B
// give u1 = 1 - U(A), u2 = U(B); the first packet is r*sin, the cached second
// packet r*cos. A fill of 8 mod 16 outputs ends with an unused cos packet.
template <int MODE>  // 0 = StdNormalGen, 1 = NormalGen(mean, sd), 2 = NormalVGen
struct NormalF32 {
    static constexpr int CALLS = 2;
    float* out;
    float mean, sd;
    const float* vmean;
    const float* vsd;
    __device__ static int64_t units(int64_t n) { return (n + 15) / 16; }
    __device__ void emit(int64_t g, int64_t per, int64_t u, int l, const uint64_t* w) const {
        float sn[2], cs[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const float u1 = 1.0f - ur_f32(h ? hi32(w[0]) : lo32(w[0]));
            const float u2 = ur_f32(h ? hi32(w[1]) : lo32(w[1]));
            const float radius = sqrtf(-2.0f * plog_pos(u1));
            float s, c;
            psincos(ERD_F(2 * kPi) * u2, s, c);
            sn[h] = radius * s;
            cs[h] = radius * c;
        }
        const int64_t base = g * per + 16 * u + 2 * l;
        const bool full = 16 * u + 16 <= per;
        if (MODE == 1) {
#pragma unroll
            for (int h = 0; h < 2; ++h) { sn[h] = sn[h] * sd + mean; cs[h] = cs[h] * sd + mean; }
        } else if (MODE == 2) {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                sn[h] = sn[h] * vsd[base + h] + vmean[base + h];
                if (full) cs[h] = cs[h] * vsd[base + 8 + h] + vmean[base + 8 + h];
            }
        }
        *reinterpret_cast<float2*>(out + base) = make_float2(sn[0], sn[1]);
        if (full) *reinterpret_cast<float2*>(out + base + 8) = make_float2(cs[0], cs[1]);
    }
};

struct ExponentialF32 {  // NormalExp.h:391-398
    static constexpr int CALLS = 1;
    float* out;
    float lambda;
    float scale;  // 1 for ExponentialGen; beta for GammaGen(alpha == 1) (NormalExp.h:561-565)
    int gamma;
    __device__ static int64_t units(int64_t n) { return n / 8; }
    __device__ void emit(int64_t g, int64_t per, int64_t u, int l, const uint64_t* w) const {
        float v[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            float x = -(plog_pos(1.0f - ur_f32(h ? hi32(w[0]) : lo32(w[0]))) / lambda);
            v[h] = gamma ? scale * x : x;
        }
        *reinterpret_cast<float2*>(out + g * per + 8 * u + 2 * l) = make_float2(v[0], v[1]);
    }
};

struct CauchyF32 {  // NormalExp.h:843-855
    static constexpr int CALLS = 1;
    float* out;
    float a, b;
    __device__ static int64_t units(int64_t n) { return n / 8; }
    __device__ void emit(int64_t g, int64_t per, int64_t u, int l, const uint64_t* w) const {
        float v[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            float s, c;
            psincos(ERD_F(kPi) * (ur_f32(h ? hi32(w[0]) : lo32(w[0])) - 0.5f), s, c);
            v[h] = a + b * (s / c);
        }
        *reinterpret_cast<float2*>(out + g * per + 8 * u + 2 * l) = make_float2(v[0], v[1]);
    }
};

struct StudentTF32 {  // NormalExp.h:285-301 (sine from the reference's psincos)
    static constexpr int CALLS = 2;
    float* out;
    float n, k;  // k = -2/n in float
    __device__ static int64_t units(int64_t n_) { return n_ / 8; }
    __device__ void emit(int64_t g, int64_t per, int64_t u, int l, const uint64_t* w) const {
        float v[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const float u1 = 1.0f - ur_f32(h ? hi32(w[0]) : lo32(w[0]));
            const float u2 = ur_f32(h ? hi32(w[1]) : lo32(w[1]));
            const float radius = sqrtf(n * (pexp(plog_pos(u1) * k) - 1.0f));
            float s, c;
            psincos(ERD_F(2 * kPi) * u2, s, c);
            v[h] = radius * s;
        }
        *reinterpret_cast<float2*>(out + g * per + 8 * u + 2 * l) = make_float2(v[0], v[1]);
    }
};

// GammaGen with integer alpha < 20: -(beta * log(prod of alpha nonzero
// uniforms)), alpha calls per packet (NormalExp.h:567-578).
template <int ALPHA>
struct GammaIntF32 {
    static constexpr int CALLS = ALPHA;
    float* out;
    float beta;
    __device__ static int64_t units(int64_t n) { return n / 8; }
    __device__ void emit(int64_t g, int64_t per, int64_t u, int l, const uint64_t* w) const {
        float v[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            float yx = nzur_f32(h ? hi32(w[0]) : lo32(w[0]));
#pragma unroll
            for (int c = 1; c < ALPHA; ++c) yx = yx * nzur_f32(h ? hi32(w[c]) : lo32(w[c]));
            v[h] = -(beta * plog(yx));
        }
        *reinterpret_cast<float2*>(out + g * per + 8 * u + 2 * l) = make_float2(v[0], v[1]);
    }
};

// --------------------------------------------------------------- kernel ---
template <class D>
__global__ __launch_bounds__(64) void k_fill_fixed(EngineState* __restrict__ states,
                                                   int64_t per, D d) {
    constexpr int C = D::CALLS;
    __shared__ uint64_t mt[kWords];
    __shared__ uint64_t carry[(C > 1 ? C - 1 : 1) * kLanes];
    const int lane = threadIdx.x;
    const int64_t g = blockIdx.x;
    EngineState* st = states + g;
    const int64_t units = D::units(per);
    if (units == 0) return;
    load_state(mt, st, lane);
    int pos = st->idx;
    int nc = 0;
    bool twisted = false;
    int64_t done = 0;
    for (;;) {
        if (pos >= kN) {
            twist_lds(mt, lane);
            pos = 0;
            twisted = true;
        }
        const int avail = nc + (kN - pos);
        const int64_t nu64 = units - done < (int64_t)(avail / C) ? units - done : (int64_t)(avail / C);
        const int nu = (int)nu64;
        for (int t = lane; t < nu * kLanes; t += 64) {
            const int u = t >> 2, l = t & 3;
            uint64_t w[C];
#pragma unroll
            for (int c = 0; c < C; ++c) {
                const int v = u * C + c;
                const uint64_t raw = (C > 1 && v < nc) ? carry[v * kLanes + l]
                                                       : mt[(pos + v - nc) * kLanes + l];
                w[c] = temper(raw);
            }
            d.emit(g, per, done + u, l, w);
        }
        const int consumed = nu * C;
        done += nu;
        if (done == units) {
            pos = pos + consumed - nc;
            break;
        }
        // Keep the words of the unit that straddles the boundary.
        // (Up to (C-1)*4 words, i.e. KEEP per thread; read all, then write.)
        constexpr int KEEP = ((C > 1 ? C - 1 : 1) * kLanes + 63) / 64;
        const int left = avail - consumed;
        uint64_t keep[KEEP];
#pragma unroll
        for (int k = 0; k < KEEP; ++k) {
            const int i = lane + 64 * k;
            keep[k] = 0;
            if (i < left * kLanes) {
                const int v = consumed + (i >> 2), l = i & 3;
                keep[k] = (v < nc) ? carry[v * kLanes + l] : mt[(pos + v - nc) * kLanes + l];
            }
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < KEEP; ++k) {
            const int i = lane + 64 * k;
            if (i < left * kLanes) carry[i] = keep[k];
        }
        __syncthreads();
        nc = left;
        pos = kN;
    }
    __syncthreads();
    if (twisted) store_state(st, mt, lane);
    if (lane == 0) st->idx = pos;
}

// ------------------------------------------------------------- launchers ---
template <class D>
static hipError_t launch_fixed(EngineState* states, int64_t n_engines, int64_t per, const D& d,
                               hipStream_t stream) {
    if (n_engines <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_fixed<D>, dim3((unsigned)n_engines), dim3(64), 0, stream, states, per, d);
    return hipGetLastError();
}

hipError_t launch_fill_u64(EngineState* st, int64_t n, uint64_t* out, int64_t per, hipStream_t s) {
    RawU64 d{out};
    return launch_fixed(st, n, per, d, s);
}

hipError_t launch_fill_uniform_f32(EngineState* st, int64_t n, float* out, int64_t per, float a,
                                   float b, int affine, hipStream_t s) {
    UniformF32 d{out, a, b - a, affine};
    return launch_fixed(st, n, per, d, s);
}

hipError_t launch_fill_uniform_f64(EngineState* st, int64_t n, double* out, int64_t per, double a,
                                   double b, int affine, hipStream_t s) {
    UniformF64 d{out, a, b - a, affine};
    return launch_fixed(st, n, per, d, s);
}

hipError_t launch_fill_normal_f32(EngineState* st, int64_t n, float* out, int64_t per, float mean,
                                  float sd, int affine, hipStream_t s) {
    if (affine) {
        NormalF32<1> d{out, mean, sd, nullptr, nullptr};
        return launch_fixed(st, n, per, d, s);
    }
    NormalF32<0> d{out, mean, sd, nullptr, nullptr};
    return launch_fixed(st, n, per, d, s);
}

hipError_t launch_fill_normal_vop_f32(EngineState* st, int64_t n, float* out, int64_t per,
                                      const float* mean, const float* sd, hipStream_t s) {
    NormalF32<2> d{out, 0.f, 1.f, mean, sd};
    return launch_fixed(st, n, per, d, s);
}

hipError_t launch_fill_exponential_f32(EngineState* st, int64_t n, float* out, int64_t per,
                                       float lambda, hipStream_t s) {
    ExponentialF32 d{out, lambda, 1.0f, 0};
    return launch_fixed(st, n, per, d, s);
}

hipError_t launch_fill_cauchy_f32(EngineState* st, int64_t n, float* out, int64_t per, float a,
                                  float b, hipStream_t s) {
    CauchyF32 d{out, a, b};
    return launch_fixed(st, n, per, d, s);
}

hipError_t launch_fill_studentt_f32(EngineState* st, int64_t n, float* out, int64_t per, float nu,
                                    hipStream_t s) {
    StudentTF32 d{out, nu, -2 / nu};
    return launch_fixed(st, n, per, d, s);
}

template <int A>
static hipError_t launch_gamma_int(EngineState* st, int64_t n, float* out, int64_t per, float beta,
                                   hipStream_t s) {
    GammaIntF32<A> d{out, beta};
    return launch_fixed(st, n, per, d, s);
}

hipError_t launch_fill_gamma_fixed_f32(EngineState* st, int64_t n, float* out, int64_t per,
                                       float alpha, float beta, hipStream_t s) {
    if (alpha == 1.0f) {
        ExponentialF32 d{out, 1.0f, beta, 1};
        return launch_fixed(st, n, per, d, s);
    }
    switch ((int)alpha) {
        case 2: return launch_gamma_int<2>(st, n, out, per, beta, s);
        case 3: return launch_gamma_int<3>(st, n, out, per, beta, s);
        case 4: return launch_gamma_int<4>(st, n, out, per, beta, s);
        case 5: return launch_gamma_int<5>(st, n, out, per, beta, s);
        case 6: return launch_gamma_int<6>(st, n, out, per, beta, s);
        case 7: return launch_gamma_int<7>(st, n, out, per, beta, s);
        case 8: return launch_gamma_int<8>(st, n, out, per, beta, s);
        case 9: return launch_gamma_int<9>(st, n, out, per, beta, s);
        case 10: return launch_gamma_int<10>(st, n, out, per, beta, s);
        case 11: return launch_gamma_int<11>(st, n, out, per, beta, s);
        case 12: return launch_gamma_int<12>(st, n, out, per, beta, s);
        case 13: return launch_gamma_int<13>(st, n, out, per, beta, s);
        case 14: return launch_gamma_int<14>(st, n, out, per, beta, s);
        case 15: return launch_gamma_int<15>(st, n, out, per, beta, s);
        case 16: return launch_gamma_int<16>(st, n, out, per, beta, s);
        case 17: return launch_gamma_int<17>(st, n, out, per, beta, s);
        case 18: return launch_gamma_int<18>(st, n, out, per, beta, s);
        case 19: return launch_gamma_int<19>(st, n, out, per, beta, s);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace erd
