// Fills whose consumption per output packet is fixed: raw u64, uniformReal
// (f32/f64), normal (Box-Muller pairs), exponential, Cauchy, Student-t, gamma
// with alpha == 1 or integer alpha < 20, and NormalVGen (per-element mean/sd).
//
// One workgroup = one wave64 = one Vmt19937_64 engine. The engine's 312x4 u64
// state is loaded into LDS once, twisted in place block by block, and written
// back only if a twist happened. A "unit" is the group of engine draws that
// produces one output packet (or a Box-Muller packet pair); each thread takes
// (unit, lane) items, so every MT word is read and tempered exactly once.
// Units that straddle a block boundary keep their first words in a small LDS
// carry area across the twist (op_fixed in erd_ops.hpp).
#include <algorithm>
#include <cmath>
#include "erd_device.hpp"
#include "erd_dists.hpp"
#include "erd_kernels.hpp"
#include "erd_ops.hpp"

namespace erd {

template <class D>
__global__ __launch_bounds__(64) void k_fill_fixed(EngineState* __restrict__ states,
                                                   int64_t per, D d) {
    constexpr int C = D::CALLS;
    __shared__ uint64_t mt[kWords];
    __shared__ uint64_t carry[(C > 1 ? C - 1 : 1) * kLanes];
    const int lane = threadIdx.x;
    const int64_t g = xcd_block(blockIdx.x, gridDim.x);
    if (D::units(per) == 0) return;
    Eng e;
    eng_load(e, mt, states + g, lane);
    d.rebase(g * per);
    op_fixed(e, carry, per, d, lane);
    eng_store(e, states + g, lane);
}

// k_fill_fixed with two-lane items and 16-byte stores (op_fixed2).
template <class D>
__global__ __launch_bounds__(64) void k_fill_fixed2(EngineState* __restrict__ states,
                                                    int64_t per, D d) {
    constexpr int C = D::CALLS;
    __shared__ uint64_t mt[kWords];
    __shared__ uint64_t carry[(C > 1 ? C - 1 : 1) * kLanes];
    const int lane = threadIdx.x;
    const int64_t g = xcd_block(blockIdx.x, gridDim.x);
    if (D::units(per) == 0) return;
    Eng e;
    eng_load(e, mt, states + g, lane);
    d.rebase(g * per);
    op_fixed2(e, carry, per, d, lane);
    eng_store(e, states + g, lane);
}

// Plain 16-byte stores of a constant, one workgroup of 64 threads per
// contiguous block of `per` floats, blocks placed on the XCDs as the fills
// place their engines: the store pattern of the fills without their
// arithmetic (the practical HBM-write ceiling they are measured against).
// per is a multiple of 4.
__global__ __launch_bounds__(64) void k_store_blocks(float* __restrict__ out, int64_t per, float v) {
    float4* o = reinterpret_cast<float4*>(out + xcd_block(blockIdx.x, gridDim.x) * per);
    const int64_t n4 = per / 4;
    const float4 x = make_float4(v, v, v, v);
    for (int64_t i = threadIdx.x; i < n4; i += 64) o[i] = x;
}

// Reused normal generator: the cached cos packet (count 8) is emitted first,
// in the call's affine form; a trailing half unit leaves its cos packet in gs.
template <int MODE>
__global__ __launch_bounds__(64) void k_fill_normal_gen(EngineState* __restrict__ states, int64_t per,
                                                        NormalF32<MODE> d, GenState* __restrict__ gs_all) {
    __shared__ uint64_t mt[kWords];
    __shared__ uint64_t carry[kLanes];
    const int lane = threadIdx.x;
    const int64_t g = blockIdx.x;
    if (per < 8) return;
    GenState* gs = gs_all + g;
    Eng e;
    eng_load(e, mt, states + g, lane);
    d.rebase(g * per);
    int64_t n = per;
    if (gs->count == 8) {
        if (lane < 8) {
            const float z = gs->v[lane];
            float x = z;
            if (MODE == 1) x = z * d.sd + d.mean;
            if (MODE == 3) x = pexp(z * d.sd + d.mean);
            if (MODE == 2) x = z * d.vs(lane) + d.vm(lane);
            d.out[lane] = x;
        }
        d.rebase(8);
        n -= 8;
    }
    __syncthreads();
    d.gs = gs;
    op_fixed(e, carry, n, d, lane);
    if (lane == 0) gs->count = (n % 16 == 8) ? 8 : 0;
    eng_store(e, states + g, lane);
}

// DiscreteGen<int32_t, int32_t>::packetOp under AVX2 (Discrete.h:443-483):
// one draw gives a Packet8i of categories, the low half now and the high half
// from the cache on the next call; int j of the draw is 32-bit word j. A fill
// of 4 mod 8 ends on a low half and drops the cached one with the generator.
struct DiscreteI32 {
    static constexpr int CALLS = 1;
    int32_t* out;
    DiscreteTable t;
    __device__ void rebase(int64_t off) { out += off; }
    __device__ static int64_t units(int64_t n) { return (n + 7) / 8; }
    __device__ void emit(int64_t per, int64_t u, int l, const uint64_t* w) const {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int64_t i = 8 * u + 2 * l + h;
            if (i < per) out[i] = discrete_packet(t, h ? hi32(w[0]) : lo32(w[0]));
        }
    }
};

// DiscreteGen<int32_t, double>::packetOp at Packet4i under AVX with Eigen 3.x
// (Discrete.h:699-779, the EIGEN_VECTORIZE_AVX branches).
// K < 16: one full draw per packet gives 4 doubles (bit_to_ur_double of the
// u64 lanes) and lane l = K - #{c in cdf : rx_l < c}.
struct DiscreteDCdfI32 {
    static constexpr int CALLS = 1;
    int32_t* out;
    const double* cdf;
    int32_t k;
    __device__ void rebase(int64_t off) { out += off; }
    __device__ static int64_t units(int64_t n) { return n / 4; }
    __device__ void emit(int64_t per, int64_t u, int l, const uint64_t* w) const {
        const double rx = ur_f64(w[0]);
        int32_t r = k;
        for (int c = 0; c < k; ++c) r -= rx < cdf[c] ? 1 : 0;
        out[4 * u + l] = r;
    }
};

// K >= 16: a packet takes idx = the 4 int32 words of half() & bitmask, then a
// full draw's 4 doubles u, and lane l is idx_l if u_l < prob[idx_l] (the
// double alias table) else alias[idx_l]. half() serves the low half of a new
// draw and caches its high half for the next packet, so with an empty cache
// two packets take three draws: the split draw, then one full draw each.
struct DiscreteDAlias {
    const double* prob;
    const int32_t* alias;
    uint32_t bitmask;
    __device__ int32_t pick(uint32_t word, uint64_t full) const {
        const uint32_t idx = word & bitmask;
        return ur_f64(full) < prob[idx] ? (int32_t)idx : alias[idx];
    }
    // int32 word l of the half made of u64 lanes (h0, h1) of the split draw,
    // whose lane l sits in thread (quad base + l) of a unit's four threads
    __device__ static uint32_t half_word(uint64_t mine, int l, int first_lane) {
        const int src = (int)(threadIdx.x & ~3u) + first_lane + (l >> 1);
        const uint32_t lo = (uint32_t)__shfl((int)(uint32_t)mine, src);
        const uint32_t hi = (uint32_t)__shfl((int)(uint32_t)(mine >> 32), src);
        return (l & 1) ? hi : lo;
    }
};

// The first packet when the engine holds a cached half (cache words c0, c1).
struct DiscreteDFirst {
    int32_t* out;
    DiscreteDAlias t;
    uint64_t c0, c1;
    __device__ static int64_t units(int64_t n) { return n / 4; }
    template <class F>
    __device__ void emit(int64_t per, int64_t u, int l, F fetch) const {
        const uint64_t h = (l >> 1) ? c1 : c0;
        out[l] = t.pick((l & 1) ? (uint32_t)(h >> 32) : (uint32_t)h, fetch(0));
    }
};

// Packet pairs: draws (split, full A, full B) per unit. The last unit's split
// draw leaves its high half in hc (the reference keeps it in the engine's
// cache member even once served).
struct DiscreteDPair {
    int32_t* out;
    DiscreteDAlias t;
    uint64_t* hc;  // LDS, 2 words
    int64_t n_units;
    __device__ static int64_t units(int64_t n) { return n / 8; }
    template <class F>
    __device__ void emit(int64_t per, int64_t u, int l, F fetch) const {
        const uint64_t s = fetch(0);
        out[8 * u + l] = t.pick(DiscreteDAlias::half_word(s, l, 0), fetch(1));
        out[8 * u + 4 + l] = t.pick(DiscreteDAlias::half_word(s, l, 2), fetch(2));
        if (u == n_units - 1 && l >= 2) hc[l - 2] = s;
    }
};

// A trailing single packet with an empty cache: (split, full); the split
// draw's high half becomes the engine's valid cache (hc).
struct DiscreteDLast {
    int32_t* out;
    DiscreteDAlias t;
    uint64_t* hc;
    __device__ static int64_t units(int64_t n) { return n / 4; }
    template <class F>
    __device__ void emit(int64_t per, int64_t u, int l, F fetch) const {
        const uint64_t s = fetch(0);
        out[l] = t.pick(DiscreteDAlias::half_word(s, l, 0), fetch(1));
        if (l >= 2) hc[l - 2] = s;
    }
};

__global__ __launch_bounds__(64) void k_fill_discrete_d_alias(EngineState* __restrict__ states,
                                                              int32_t* __restrict__ out, int64_t per,
                                                              DiscreteDAlias t) {
    __shared__ uint64_t mt[kWords];
    __shared__ uint64_t carry[2 * kLanes];
    __shared__ uint64_t hc[2];
    const int lane = threadIdx.x;
    const int64_t g = blockIdx.x;
    const int64_t T = per / 4;
    if (T == 0) return;
    Eng e;
    eng_load(e, mt, states + g, lane);
    int32_t* o = out + g * per;
    int64_t k0 = 0;
    if (e.half_valid) {
        op_fixed_rt<2>(e, carry, 4, 1, DiscreteDFirst{o, t, e.half_cache[0], e.half_cache[1]}, lane);
        e.half_valid = 0;
        k0 = 1;
    }
    const int64_t pairs = (T - k0) / 2;
    if (pairs > 0) {
        op_fixed_rt<2>(e, carry, 8 * pairs, 3, DiscreteDPair{o + 4 * k0, t, hc, pairs}, lane);
        e.half_cache[0] = hc[0];  // op_fixed_rt ends with a barrier
        e.half_cache[1] = hc[1];
    }
    if ((T - k0) & 1) {
        op_fixed_rt<2>(e, carry, 4, 2, DiscreteDLast{o + 4 * (k0 + 2 * pairs), t, hc}, lane);
        e.half_cache[0] = hc[0];
        e.half_cache[1] = hc[1];
        e.half_valid = 1;
    }
    eng_store(e, states + g, lane);
}

// StdNormalGen fills of a list of engines, one workgroup of NT threads per
// engine (job), so that one long fill on one engine (a multivariate
// instance's Z) is not held to one wave's rate. Outputs and engine states
// equal k_fill_fixed<NormalF32<0>> on each engine. A block of 312 draws holds
// 624 items (4 normals each): at 256 threads the busiest thread takes 3 per
// block, at 640 threads (10 waves) one, which shortens a long serial fill
// when there are few engines (the Bartlett normals of a Wishart batch).
template <int NT>
__global__ __launch_bounds__(NT) void k_std_normal_jobs(EngineState* __restrict__ states,
                                                        const FillJob* __restrict__ jobs) {
    __shared__ uint64_t mt[kWords];
    __shared__ uint64_t carry[kLanes];
    const int tid = threadIdx.x;
    const FillJob jb = jobs[blockIdx.x];
    if (NormalF32<0>::units(jb.count) == 0) return;
    const int wl = tid < 64 ? tid : (1 << 20);  // state copies by wave 0
    Eng e;
    eng_load(e, mt, states + jb.engine, wl);
    NormalF32<0> d{jb.out, 0.f, 1.f, nullptr, nullptr};
    op_fixed_wg<NT>(e, carry, jb.count, d, tid);
    eng_store(e, states + jb.engine, wl);
}

hipError_t launch_std_normal_jobs(EngineState* st, const FillJob* jobs, int n_jobs, hipStream_t s, bool wide) {
    if (n_jobs <= 0) return hipSuccess;
    if (wide)
        hipLaunchKernelGGL(k_std_normal_jobs<kJobThreadsWide>, dim3((unsigned)n_jobs), dim3(kJobThreadsWide), 0, s,
                           st, jobs);
    else
        hipLaunchKernelGGL(k_std_normal_jobs<kJobThreads>, dim3((unsigned)n_jobs), dim3(kJobThreads), 0, s, st,
                           jobs);
    return hipGetLastError();
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

hipError_t launch_fill_layout(EngineState* st, int64_t n, int W, int kind, void* out, int64_t per, double a,
                              double b, int mode, hipStream_t s) {
    const float fa = (float)a, fb = (float)b;
    if (W == 4) {
        switch (kind) {
            case LK_U64: return launch_fixed(st, n, per, ElemU64W4<RawU64>{RawU64{(uint64_t*)out}}, s);
            case LK_UNIFORM_F32:
                return launch_fixed(st, n, per, ElemF32W4<UniformF32>{UniformF32{(float*)out, fa, fb - fa, mode}}, s);
            case LK_UNIFORM_F64:
                return launch_fixed(st, n, per, ElemU64W4<UniformF64>{UniformF64{(double*)out, a, b - a, mode}}, s);
            case LK_EXPONENTIAL_F32:
                return launch_fixed(st, n, per, ElemF32W4<ExponentialF32>{ExponentialF32{(float*)out, fa, 1.0f, 0}}, s);
            case LK_CAUCHY_F32: return launch_fixed(st, n, per, ElemF32W4<CauchyF32>{CauchyF32{(float*)out, fa, fb}}, s);
            case LK_WEIBULL_F32:
                return launch_fixed(st, n, per, ElemF32W4<WeibullF32>{WeibullF32{(float*)out, 1 / fa, fb}}, s);
            case LK_EXTREME_VALUE_F32:
                return launch_fixed(st, n, per, ElemF32W4<ExtremeValueF32>{ExtremeValueF32{(float*)out, fa, fb}}, s);
            case LK_EXPONENTIAL_F64:
                return launch_fixed(st, n, per, ElemU64W4<ExponentialF64>{ExponentialF64{(double*)out, a}}, s);
            case LK_CAUCHY_F64: return launch_fixed(st, n, per, ElemU64W4<CauchyF64>{CauchyF64{(double*)out, a, b}}, s);
            case LK_NORMAL_F32:
                if (mode == 1) return launch_fixed(st, n, per, NormalF32L<4, 1>{(float*)out, fa, fb}, s);
                if (mode == 3) return launch_fixed(st, n, per, NormalF32L<4, 3>{(float*)out, fa, fb}, s);
                return launch_fixed(st, n, per, NormalF32L<4, 0>{(float*)out, fa, fb}, s);
            case LK_STUDENT_T_F32: return launch_fixed(st, n, per, StudentTF32L<4>{(float*)out, fa, -2 / fa}, s);
            case LK_NORMAL_F64:
                if (mode == 1) return launch_fixed(st, n, per, NormalF64L<4, 1>{(double*)out, a, b}, s);
                return launch_fixed(st, n, per, NormalF64L<4, 0>{(double*)out, a, b}, s);
        }
    } else if (W == 16) {
        switch (kind) {
            case LK_NORMAL_F32:
                if (mode == 1) return launch_fixed(st, n, per, NormalF32L<16, 1>{(float*)out, fa, fb}, s);
                if (mode == 3) return launch_fixed(st, n, per, NormalF32L<16, 3>{(float*)out, fa, fb}, s);
                return launch_fixed(st, n, per, NormalF32L<16, 0>{(float*)out, fa, fb}, s);
            case LK_STUDENT_T_F32: return launch_fixed(st, n, per, StudentTF32L<16>{(float*)out, fa, -2 / fa}, s);
            case LK_NORMAL_F64:
                if (mode == 1) return launch_fixed(st, n, per, NormalF64L<16, 1>{(double*)out, a, b}, s);
                return launch_fixed(st, n, per, NormalF64L<16, 0>{(double*)out, a, b}, s);
        }
    }
    return hipErrorInvalidValue;
}

// Whole-refill fills drop the scalars left in an adaptor's operator() buffer:
// the buffer position (record 2k's `reserved`) goes back to 0.
__global__ void k_p8_clear_position(EngineState* __restrict__ states, int64_t n_p8) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n_p8) states[2 * k].reserved = 0;
}

hipError_t launch_fill_p8(EngineState* st, int64_t n, void* out, int64_t per_v, int mode,
                          hipStream_t s) {
    hipError_t rc;
    switch (mode) {
        case 0: rc = launch_fixed(st, n, per_v, P8Fill<0>{out, per_v, 0}, s); break;
        case 1: rc = launch_fixed(st, n, per_v, P8Fill<1>{out, per_v, 0}, s); break;
        default: rc = launch_fixed(st, n, per_v, P8Fill<2>{out, per_v, 0}, s); break;
    }
    if (rc != hipSuccess || n < 2) return rc;
    const int64_t n_p8 = n / 2;
    hipLaunchKernelGGL(k_p8_clear_position, dim3((unsigned)((n_p8 + 255) / 256)), dim3(256), 0, s, st, n_p8);
    return hipGetLastError();
}

template <class D>
static hipError_t launch_fixed2(EngineState* states, int64_t n_engines, int64_t per, const D& d,
                                hipStream_t stream) {
    if (n_engines <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_fixed2<D>, dim3((unsigned)n_engines), dim3(64), 0, stream, states, per, d);
    return hipGetLastError();
}

hipError_t launch_fill_uniform_f32(EngineState* st, int64_t n, float* out, int64_t per, float a,
                                   float b, int affine, hipStream_t s) {
    UniformF32 d{out, a, b - a, affine};
    return launch_fixed2(st, n, per, d, s);
}

hipError_t launch_fill_normal_f32_fast(EngineState* st, int64_t n, float* out, int64_t per, float mean,
                                       float sd, int affine, hipStream_t s) {
    if (affine) return launch_fixed2(st, n, per, NormalF32Fast<1>{out, mean, sd}, s);
    return launch_fixed2(st, n, per, NormalF32Fast<0>{out, mean, sd}, s);
}

hipError_t launch_store_blocks(float* out, int64_t n_blocks, int64_t per, hipStream_t s) {
    if (n_blocks <= 0 || per <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_store_blocks, dim3((unsigned)n_blocks), dim3(64), 0, s, out, per, 1.0f);
    return hipGetLastError();
}

hipError_t launch_fill_uniform_f64(EngineState* st, int64_t n, double* out, int64_t per, double a,
                                   double b, int affine, hipStream_t s) {
    UniformF64 d{out, a, b - a, affine};
    return launch_fixed(st, n, per, d, s);
}

hipError_t launch_fill_discrete_i32(EngineState* st, int64_t n, int32_t* out, int64_t per, const uint32_t* cdf,
                                   const int32_t* prob, const int32_t* alias, int32_t ncdf, uint32_t bitmask,
                                   hipStream_t s) {
    return launch_fixed(st, n, per, DiscreteI32{out, DiscreteTable{cdf, prob, alias, ncdf, bitmask}}, s);
}

hipError_t launch_fill_discrete_d_i32(EngineState* st, int64_t n, int32_t* out, int64_t per, const double* cdf,
                                      int32_t ncdf, const double* prob, const int32_t* alias, uint32_t bitmask,
                                      hipStream_t s) {
    if (n <= 0) return hipSuccess;
    if (cdf) return launch_fixed(st, n, per, DiscreteDCdfI32{out, cdf, ncdf}, s);
    hipLaunchKernelGGL(k_fill_discrete_d_alias, dim3((unsigned)n), dim3(64), 0, s, st, out, per,
                       DiscreteDAlias{prob, alias, bitmask});
    return hipGetLastError();
}

hipError_t launch_fill_normal_f64(EngineState* st, int64_t n, double* out, int64_t per, double mean,
                                  double sd, int affine, hipStream_t s) {
    if (affine) return launch_fixed(st, n, per, NormalF64<1>{out, mean, sd}, s);
    return launch_fixed(st, n, per, NormalF64<0>{out, mean, sd}, s);
}

hipError_t launch_fill_exponential_f64(EngineState* st, int64_t n, double* out, int64_t per,
                                       double lambda, hipStream_t s) {
    return launch_fixed(st, n, per, ExponentialF64{out, lambda}, s);
}

hipError_t launch_fill_cauchy_f64(EngineState* st, int64_t n, double* out, int64_t per, double a,
                                  double b, hipStream_t s) {
    return launch_fixed(st, n, per, CauchyF64{out, a, b}, s);
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
                                      const float* mean, const float* sd, hipStream_t s, float mean0,
                                      float sd0) {
    NormalF32<2> d{out, mean0, sd0, mean, sd};
    return launch_fixed(st, n, per, d, s);
}

hipError_t launch_fill_normal_gen_f32(EngineState* st, int64_t n, float* out, int64_t per, int mode,
                                      float mean, float sd, const float* vmean, const float* vsd,
                                      GenState* gs, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    if (mode == 0) {
        hipLaunchKernelGGL(k_fill_normal_gen<0>, dim3((unsigned)n), dim3(64), 0, s, st, per,
                           NormalF32<0>{out, mean, sd, nullptr, nullptr, nullptr}, gs);
    } else if (mode == 1) {
        hipLaunchKernelGGL(k_fill_normal_gen<1>, dim3((unsigned)n), dim3(64), 0, s, st, per,
                           NormalF32<1>{out, mean, sd, nullptr, nullptr, nullptr}, gs);
    } else if (mode == 3) {
        hipLaunchKernelGGL(k_fill_normal_gen<3>, dim3((unsigned)n), dim3(64), 0, s, st, per,
                           NormalF32<3>{out, mean, sd, nullptr, nullptr, nullptr}, gs);
    } else {
        hipLaunchKernelGGL(k_fill_normal_gen<2>, dim3((unsigned)n), dim3(64), 0, s, st, per,
                           NormalF32<2>{out, 0.f, 1.f, vmean, vsd, nullptr}, gs);
    }
    return hipGetLastError();
}

hipError_t launch_fill_lognormal_f32(EngineState* st, int64_t n, float* out, int64_t per, float mean,
                                     float sd, hipStream_t s) {
    NormalF32<3> d{out, mean, sd, nullptr, nullptr, nullptr};
    return launch_fixed(st, n, per, d, s);
}

hipError_t launch_fill_weibull_f32(EngineState* st, int64_t n, float* out, int64_t per, float a, float b,
                                   hipStream_t s) {
    WeibullF32 d{out, 1 / a, b};
    return launch_fixed(st, n, per, d, s);
}

hipError_t launch_fill_extreme_value_f32(EngineState* st, int64_t n, float* out, int64_t per, float a,
                                         float b, hipStream_t s) {
    ExtremeValueF32 d{out, a, b};
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

// ------------------------------------------- raw-bit and VoP float fills ---
hipError_t launch_fill_bernoulli_f32(EngineState* st, int64_t n, float* out, int64_t per, int32_t p,
                                     const float* vp, hipStream_t s) {
    if (vp) {
        BernoulliF32<1> d{out, 0, vp};
        return launch_fixed(st, n, per, d, s);
    }
    BernoulliF32<0> d{out, p, nullptr};
    return launch_fixed(st, n, per, d, s);
}

hipError_t launch_fill_balanced_f32(EngineState* st, int64_t n, float* out, int64_t per, float a, float b,
                                    const float* va, const float* vb, hipStream_t s) {
    if (va) {
        BalancedF32<1> d{out, 0.f, 0.f, va, vb};
        return launch_fixed(st, n, per, d, s);
    }
    BalancedF32<0> d{out, b - a, a, nullptr, nullptr};
    return launch_fixed(st, n, per, d, s);
}

hipError_t launch_fill_vop_f32(EngineState* st, int64_t n, int kind, float* out, int64_t per, const float* va,
                               const float* vb, hipStream_t s) {
    switch (kind) {
        case VOP_UNIFORM: return launch_fixed(st, n, per, VopF32<0>{out, va, vb}, s);
        case VOP_EXPONENTIAL: return launch_fixed(st, n, per, VopF32<1>{out, va, nullptr}, s);
        case VOP_CAUCHY: return launch_fixed(st, n, per, VopF32<2>{out, va, vb}, s);
        case VOP_WEIBULL: return launch_fixed(st, n, per, VopF32<3>{out, va, vb}, s);
        case VOP_EXTREME_VALUE: return launch_fixed(st, n, per, VopF32<4>{out, va, vb}, s);
        case VOP_STUDENT_T: return launch_fixed(st, n, per, StudentTVopF32{out, va}, s);
        case VOP_LOGNORMAL:
            return launch_fixed(st, n, per, NormalF32<4>{out, 0.f, 1.f, va, vb, nullptr}, s);
        default: return hipErrorInvalidValue;
    }
}

// ----------------------------------------------------- half-draw int32 fills ---
template <class D>
__global__ __launch_bounds__(64) void k_fill_half(EngineState* __restrict__ states, int32_t* __restrict__ out,
                                                  int64_t per, D d) {
    __shared__ uint64_t mt[kWords];
    __shared__ uint64_t carry[2];
    const int lane = threadIdx.x;
    const int64_t g = blockIdx.x;
    if (per / 4 == 0) return;
    Eng e;
    eng_load(e, mt, states + g, lane);
    d.out = out + g * per;
    op_half_fixed(e, carry, per / 4, d, lane);
    eng_store(e, states + g, lane);
}

template <class D>
static hipError_t launch_half(EngineState* st, int64_t n, int32_t* out, int64_t per, const D& d, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_half<D>, dim3((unsigned)n), dim3(64), 0, s, st, out, per, d);
    return hipGetLastError();
}

hipError_t launch_fill_bernoulli_i32(EngineState* st, int64_t n, int32_t* out, int64_t per, int32_t p,
                                     hipStream_t s) {
    return launch_half(st, n, out, per, BernoulliI32{nullptr, p}, s);
}

hipError_t launch_fill_geometric_i32(EngineState* st, int64_t n, int32_t* out, int64_t per, float rlog_q,
                                     hipStream_t s) {
    return launch_half(st, n, out, per, GeometricI32{nullptr, rlog_q}, s);
}

hipError_t launch_fill_discrete_f_i32(EngineState* st, int64_t n, int32_t* out, int64_t per, const float* tab,
                                      const int32_t* alias, int32_t k, uint32_t bitmask, hipStream_t s) {
    if (!alias) return launch_half(st, n, out, per, DiscreteFI32<1>{nullptr, tab, nullptr, k, 0u}, s);
    return launch_half(st, n, out, per, DiscreteFI32<2>{nullptr, tab, alias, 0, bitmask}, s);
}

hipError_t launch_fill_randbits_i32(EngineState* st, int64_t n, int32_t* out, int64_t per, hipStream_t s) {
    return launch_half(st, n, out, per, RandbitsI32{nullptr}, s);
}

// UniformIntGen<int32_t>: a constant range writes min without drawing; a
// power-of-two range masks one half draw per packet; otherwise the
// CompressMask chain of op_uniform_int.
__global__ __launch_bounds__(64) void k_fill_uniform_int(EngineState* __restrict__ states, int32_t* __restrict__ out,
                                                         int64_t per, UniformIntParams u) {
    __shared__ uint64_t mt[kWords];
    __shared__ uint64_t carry[2];
    __shared__ int32_t ring[kUiRing];
    const int lane = threadIdx.x;
    const int64_t g = blockIdx.x;
    if (per / 4 == 0) return;
    Eng e;
    eng_load(e, mt, states + g, lane);
    op_uniform_int(e, carry, ring, out + g * per, per / 4, u, lane);
    eng_store(e, states + g, lane);
}

__global__ void k_fill_const_i32(int32_t* __restrict__ out, int64_t n, int32_t v) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = v;
}

hipError_t launch_fill_uniform_int_i32(EngineState* st, int64_t n, int32_t* out, int64_t per, int32_t lo,
                                       uint32_t pdiff, hipStream_t s) {
    if (n <= 0 || per <= 0) return hipSuccess;
    if (pdiff == 0) {
        const int64_t total = n * per;
        const int64_t blocks = std::min<int64_t>((total + 255) / 256, 65536);
        hipLaunchKernelGGL(k_fill_const_i32, dim3((unsigned)blocks), dim3(256), 0, s, out, total, lo);
        return hipGetLastError();
    }
    // UniformIntGen's constructor (Discrete.h:229-246) for pdiff < 2^31
    const int bits = (int)std::ceil(std::log2((double)pdiff + 1.0));
    const uint32_t mask = (uint32_t)((1ull << bits) - 1);
    if (pdiff == mask) return launch_half(st, n, out, per, UniformIntPow2I32{nullptr, mask, lo}, s);
    int cands = 1;
    for (int bitcnt = bits; bitcnt + bits < 32; bitcnt += bits) ++cands;
    const UniformIntParams u{mask, (int32_t)(pdiff + 1), lo, bits, cands};
    hipLaunchKernelGGL(k_fill_uniform_int, dim3((unsigned)n), dim3(64), 0, s, st, out, per, u);
    return hipGetLastError();
}

}  // namespace erd
