// Packet-lockstep samplers on half draws (Packet4f through half(),
// PacketRandomEngine.h:325-336): a packet of 4 int32 consumes a data-dependent
// number of slots, so packets are produced one after another per engine.
//
//  - PoissonGen, mean < 12 (Discrete.h:860-873): Knuth's product loop runs
//    until no lane's product exceeds e^-mean; the packet consumes
//    1 + max(lane counts) half draws. Also BinomialGen with n >= 25 and
//    n*min(p,1-p) < 1 (Discrete.h:1007-1010), plus its final flip.
//  - BinomialGen, n < 25 (Discrete.h:996-1006): n half draws per packet, each
//    lane counts draws >= p. Slots are independent, so up to 16 are compared
//    at once by the whole wave and counted with ballots.
//  - BinomialVGen<int32_t> (Discrete.h:1124-1192): per-lane parameters; the
//    lanes with mean < 1 run the Knuth loop in lockstep, then the others run
//    tangent rejection keeping each lane's first accepted candidate. Sixteen
//    candidate iterations are evaluated at once (thread 4i+j runs iteration i
//    of lane j); a ballot per lane finds its first acceptance, and the packet
//    consumes exactly the iterations up to the slowest lane, as the sequential
//    loop would.
#include <cmath>
#include "erd_device.hpp"
#include "erd_kernels.hpp"
#include "erd_slots.hpp"

namespace erd {

constexpr int kHalfSlots = kWords / 2;   // 624 half draws per block
constexpr int kLockstepCap = 1 << 20;    // iterations per packet before giving up

// Half-draw cursor over the engine in LDS (at most one carried slot).
struct HalfCursor {
    uint64_t* mt;
    uint64_t* carry;
    int pos;
    int nc;
    bool twisted;
    bool used;  // a block slot was consumed by this fill

    __device__ int avail() const { return nc + (kHalfSlots - pos); }
    // Tempered 32-bit word of float lane j (0..3) in virtual slot v.
    __device__ uint32_t word32(int v, int j) const {
        const uint64_t w = slot_word<2>(mt, carry, nc, pos, v, j >> 1);
        return (j & 1) ? hi32(w) : lo32(w);
    }
    __device__ void advance(int n) {
        if (n >= nc) {
            if (n > nc) used = true;
            pos += n - nc;
            nc = 0;
        }
    }
    // Make at least `need` (<= 2) slots available: keep what is left of the
    // block in the carry area and twist. Called by the whole wave.
    __device__ void ensure(int need, int lane) {
        const int a = avail();
        if (a >= need) return;
        carry_tail<2, 1>(mt, carry, nc, pos, 0, a, lane);
        nc = a;
        twist_lds(mt, lane);
        pos = 0;
        twisted = true;
    }
};

__device__ __forceinline__ void cursor_begin(HalfCursor& cur, uint64_t* mt, uint64_t* carry,
                                             const EngineState* st, int lane) {
    load_state(mt, st, lane);
    cur.mt = mt;
    cur.carry = carry;
    cur.pos = 2 * st->idx;
    cur.nc = 0;
    cur.twisted = false;
    cur.used = false;
    if (st->half_valid) {
        if (lane < 2) carry[lane] = st->half_cache[lane];
        cur.nc = 1;
    }
    __syncthreads();
}

__device__ __forceinline__ void cursor_end(const HalfCursor& cur, EngineState* st, int lane) {
    __syncthreads();
    if (cur.twisted) store_state(st, cur.mt, lane);
    if (lane == 0) {
        if (cur.nc > 0) {
            // Nothing consumed beyond the carried half: the engine keeps it.
            st->idx = cur.pos / 2;
        } else if (cur.pos & 1) {
            st->idx = (cur.pos + 1) / 2;
            st->half_valid = 1;
            st->half_cache[0] = temper(cur.mt[2 * cur.pos]);
            st->half_cache[1] = temper(cur.mt[2 * cur.pos + 1]);
        } else {
            st->idx = cur.pos / 2;
            st->half_valid = 0;
            // The reference keeps the last split draw's high half in its cache
            // member after handing it out; mirror it so records compare equal.
            if (cur.used) {
                st->half_cache[0] = temper(cur.mt[2 * (cur.pos - 1)]);
                st->half_cache[1] = temper(cur.mt[2 * (cur.pos - 1) + 1]);
            }
        }
    }
}

// Knuth loop of PoissonGen::packetOp (Discrete.h:860-873) for lane j = lane&3;
// returns the lane's count. Wave-uniform control flow.
__device__ __forceinline__ int32_t knuth_packet(HalfCursor& cur, float thr, int lane,
                                                uint32_t* status) {
    const int j = lane & 3;
    float val = 1.0f;
    int32_t res = 0;
    for (int it = 0;; ++it) {
        cur.ensure(1, lane);
        const float u = ur_f32(cur.word32(0, j));
        cur.advance(1);
        val = val * u;
        const bool c = thr < val;
        const uint64_t any = __ballot(c) & 0xFull;
        if (!any) break;
        res += c ? 1 : 0;
        if (it > kLockstepCap) {
            if (lane == 0) atomicOr(status, kStatusIterCap);
            break;
        }
    }
    return res;
}

// PoissonGen (mean < 12) and BinomialGen's Poisson regime (flip applied).
__global__ __launch_bounds__(64) void k_poisson_knuth(EngineState* __restrict__ states,
                                                      int32_t* __restrict__ out, int64_t per,
                                                      float thr, int32_t trials, int flip,
                                                      uint32_t* __restrict__ status) {
    __shared__ uint64_t mt[kWords];
    __shared__ uint64_t carry[2];
    const int lane = threadIdx.x;
    EngineState* st = states + blockIdx.x;
    int32_t* o = out + (int64_t)blockIdx.x * per;
    const int64_t T = per / 4;
    if (T == 0) return;
    HalfCursor cur;
    cursor_begin(cur, mt, carry, st, lane);
    for (int64_t pk = 0; pk < T; ++pk) {
        const int32_t r = knuth_packet(cur, thr, lane, status);
        if (lane < 4) o[4 * pk + lane] = flip ? trials - r : r;
    }
    cursor_end(cur, st, lane);
}

// BinomialGen, trials < 25: res = n - #{draws >= p}.
__global__ __launch_bounds__(64) void k_binomial_small(EngineState* __restrict__ states,
                                                       int32_t* __restrict__ out, int64_t per,
                                                       int32_t trials, float pp) {
    __shared__ uint64_t mt[kWords];
    __shared__ uint64_t carry[2];
    const int lane = threadIdx.x;
    const int i = lane >> 2, j = lane & 3;
    EngineState* st = states + blockIdx.x;
    int32_t* o = out + (int64_t)blockIdx.x * per;
    const int64_t T = per / 4;
    if (T == 0) return;
    HalfCursor cur;
    cursor_begin(cur, mt, carry, st, lane);
    const uint64_t jmask = 0x1111111111111111ull << j;
    for (int64_t pk = 0; pk < T; ++pk) {
        int32_t res = trials;
        int left = trials;
        while (left > 0) {
            cur.ensure(1, lane);
            int take = left < 16 ? left : 16;
            const int a = cur.avail();
            take = take < a ? take : a;
            bool c = false;
            if (i < take) c = pp <= ur_f32(cur.word32(i, j));
            res -= __popcll(__ballot(c) & jmask);
            cur.advance(take);
            left -= take;
        }
        if (lane < 4) o[4 * pk + lane] = res;
    }
    cursor_end(cur, st, lane);
}

// BinomialVGen<int32_t>::packetOp, per-lane float parameters.
__global__ __launch_bounds__(64) void k_binomial_vop(EngineState* __restrict__ states,
                                                     int32_t* __restrict__ out, int64_t per,
                                                     const int32_t* __restrict__ trials,
                                                     const float* __restrict__ prob,
                                                     uint32_t* __restrict__ status) {
    __shared__ uint64_t mt[kWords];
    __shared__ uint64_t carry[2];
    __shared__ float res_lds[4];
    const int lane = threadIdx.x;
    const int i = lane >> 2, j = lane & 3;
    const int64_t base = (int64_t)blockIdx.x * per;
    EngineState* st = states + blockIdx.x;
    const int64_t T = per / 4;
    if (T == 0) return;
    HalfCursor cur;
    cursor_begin(cur, mt, carry, st, lane);
    const uint64_t jmask = 0x1111111111111111ull << j;
    for (int64_t pk = 0; pk < T; ++pk) {
        const int32_t n = trials[base + 4 * pk + j];
        const float p = prob[base + 4 * pk + j];
        const float psmall = (p < 1.0f - p) ? p : 1.0f - p;  // pmin(p, 1-p)
        const float ptr = (float)n;
        const float pmean = ptr * psmall;
        const float pg1 = plgamma_approx(ptr + 1.0f);
        const float psq = sqrtf((2.0f * pmean) * (1.0f - psmall));
        const float plp = plog(psmall);
        const float plq = plog(1.0f - psmall);
        bool valid = pmean < 1.0f;
        float valid_res = 0.0f;
        const uint64_t V = __ballot(valid) & 0xFull;
        if (V) {
            // Knuth loop over the lanes with mean < 1 (the others are masked).
            const float nem = pexp(-pmean);
            float val = 1.0f;
            int32_t res = 0;
            for (int it = 0;; ++it) {
                cur.ensure(1, lane);
                const float u = ur_f32(cur.word32(0, j));
                cur.advance(1);
                val = val * u;
                const bool c = (nem < val) && valid;
                if (!(__ballot(c) & 0xFull)) break;
                res += c ? 1 : 0;
                if (it > kLockstepCap) {
                    if (lane == 0) atomicOr(status, kStatusIterCap);
                    break;
                }
            }
            valid_res = (float)res;
        }
        if (V != 0xFull) {
            uint64_t done = V;  // bit j: lane j holds its value
            for (int it = 0;; it += 16) {
                cur.ensure(2, lane);
                const int ni_avail = cur.avail() / 2;
                const int ni = ni_avail < 16 ? ni_avail : 16;
                float fres = 0.0f;
                bool fv = false;
                if (i < ni) {
                    float s, c;
                    psincos(ERD_F(kPi) * ur_f32(cur.word32(2 * i, j)), s, c);
                    const float ys = s / c;
                    fres = truncf(psq * ys + pmean);
                    const float pp1 = (ERD_F(1.2) * psq) * (1.0f + ys * ys);
                    const float pp2 = pexp((((pg1 - plgamma_approx(fres + 1.0f)) -
                                             plgamma_approx((ptr + 1.0f) - fres)) +
                                            fres * plp) +
                                           (ptr - fres) * plq);
                    const bool c1 = (0.0f <= fres) && (fres <= ptr);
                    const float u2 = ur_f32(cur.word32(2 * i + 1, j));
                    fv = c1 && (u2 <= pp1 * pp2);
                }
                const uint64_t A = __ballot(fv);
                // First accepting iteration of each lane still open.
                int need = 0;  // iterations consumed by this step
                bool all = true;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    if (done & (1ull << q)) continue;
                    const uint64_t m = A & (0x1111111111111111ull << q);
                    if (m) {
                        const int first = (__ffsll((long long)m) - 1) >> 2;
                        need = need > first + 1 ? need : first + 1;
                    } else {
                        all = false;
                    }
                }
                if (!all) need = ni;
                // Lanes accepting in this step take their first accepted value.
                const uint64_t mine = A & jmask;
                const bool open = !(done & (1ull << j));
                if (open && mine) {
                    const int first = (__ffsll((long long)mine) - 1) >> 2;
                    if (i == first) res_lds[j] = fres;
                }
                __syncthreads();
                if (open && mine) valid_res = res_lds[j];
                __syncthreads();
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    if (A & (0x1111111111111111ull << q)) done |= 1ull << q;
                cur.advance(2 * need);
                if (all) break;
                if (it > kLockstepCap) {
                    if (lane == 0) atomicOr(status, kStatusIterCap);
                    break;
                }
            }
        }
        const int32_t k = (int32_t)valid_res;
        if (lane < 4) out[base + 4 * pk + j] = (p == psmall) ? k : n - k;
    }
    cursor_end(cur, st, lane);
}

hipError_t launch_fill_poisson_knuth_i32(EngineState* st, int64_t n, int32_t* out, int64_t per,
                                         double mean, int32_t trials, int flip, uint32_t* status,
                                         hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const float thr = (float)std::exp(-mean);
    hipLaunchKernelGGL(k_poisson_knuth, dim3((unsigned)n), dim3(64), 0, s, st, out, per, thr, trials,
                       flip, status);
    return hipGetLastError();
}

hipError_t launch_fill_binomial_small_i32(EngineState* st, int64_t n, int32_t* out, int64_t per,
                                          int32_t trials, double p, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_binomial_small, dim3((unsigned)n), dim3(64), 0, s, st, out, per, trials,
                       (float)p);
    return hipGetLastError();
}

hipError_t launch_fill_binomial_vop_i32(EngineState* st, int64_t n, int32_t* out, int64_t per,
                                        const int32_t* trials, const float* p, uint32_t* status,
                                        hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_binomial_vop, dim3((unsigned)n), dim3(64), 0, s, st, out, per, trials, p,
                       status);
    return hipGetLastError();
}

}  // namespace erd
