// Sampling ops on an engine resident in LDS.
//
// A kernel loads one Vmt19937_64 record into LDS (Eng), runs one or more ops
// on it, and writes it back. Ops follow the reference's consumption exactly:
//   op_fixed    - a fixed number of full draws per output unit
//   op_reject   - candidate packets through the CompressMask queue
//   op_knuth / op_binomial_small / op_binomial_vop - packet-lockstep loops
// Full-draw ops use Eng::idx only; half-draw ops (Packet4f via half()) also
// read and update the engine's cached half draw, as the reference does.
// Every op is called by all 64 threads of the wave with wave-uniform
// arguments; `lane` is the thread index.
#pragma once
#include "erd_device.hpp"
#include "erd_slots.hpp"

namespace erd {

struct P8Feed;

struct Eng {
    uint64_t* mt;            // LDS, kWords
    int idx;                 // next draw in the block (kN: twist first)
    int half_valid;
    uint64_t half_cache[2];  // tempered upper half of the last split draw
    bool twisted;            // LDS block differs from the HBM record
    P8Feed* feed;            // non-NULL: the block is refilled from a scalar engine
};

// A P8_mt19937_64 (ParallelRandomEngineAdaptor<uint64_t, Vmt19937_64, 8>,
// PacketRandomEngine.h:429-517) used as a scalar full-bit engine by packet
// distributions. Its scalar stream is out[8r + j] = mt19937_64(seed + j)[r]:
// refill r is one draw of Vmt (seed) then one of Vmt (seed + 4). The packet
// code packs scalars into packets with RawbitsMaker<..., scalar_fullbit>; the
// argument evaluation order of _mm*_set*_epi64x(rng(), ...) is unspecified in
// C++, and this follows Clang's left-to-right order (the reference's default
// toolchain preset, CMakePresets.json:16-17):
//  kP8Raw: operator() itself, 4 scalars per draw in order;
//  kP8F34: Packet8f uniform_real (rawbits_34, arch/AVX/RandUtils.h:75-131):
//          scalars a, b, c give the 32-bit words [a.lo, a.hi, b.lo, X1,
//          b.hi, c.lo, c.hi, X2], X1 = the top bytes of a.lo, a.hi, b.lo,
//          b.lo and X2 those of b.hi, c.lo, c.hi, c.hi;
//  kP8H2:  Packet4i rawbits (arch/SSE/RandUtils.h:24-41): _mm_set_epi64x of
//          two scalars, so u64 lane 0 is the second scalar, lane 1 the first.
// The engine block in LDS is filled with such packed draws (stored
// untempered), so the ordinary ops consume them; `refills` counts the blocks.
constexpr int kP8Raw = 0, kP8F34 = 1, kP8H2 = 2;
struct P8Feed {
    uint64_t* v0;   // LDS state of Vmt (seed)
    uint64_t* v1;   // LDS state of Vmt (seed + 4)
    int idx;        // draws taken from both in their block (kN: twist first)
    int b;          // scalars of refill idx - 1 already consumed (0: none pending)
    int mode;
    int refills;
};

// Scalars per engine block of packed draws.
__device__ __forceinline__ int p8_block_scalars(int mode) {
    return mode == kP8F34 ? 3 * kN : kWords;
}

// Scalar stream position (scalars into the current block of both Vmts) and
// back: refill r of the block is scalars [8r, 8r + 8).
__device__ __forceinline__ int p8_offset(int idx, int b) { return b > 0 ? 8 * (idx - 1) + b : 8 * idx; }

// Refill the packed block `mt` with the next p8_block_scalars(mode) scalars
// of the feed (twisting both Vmts once if the block runs into their next one).
__device__ __attribute__((noinline)) void p8_refill(P8Feed& f, uint64_t* mt, int lane) {
    const int n = p8_block_scalars(f.mode);
    const int o0 = p8_offset(f.idx, f.b);
    constexpr int kBlock = 8 * kN;  // scalars per block of both Vmts
    // scalars still in the current block of the Vmts
    for (int q = lane; q < n; q += 64) {
        const int o = o0 + q;
        if (o < kBlock) {
            const uint64_t* v = ((o & 7) < 4) ? f.v0 : f.v1;
            mt[q] = temper(v[(o >> 3) * kLanes + (o & 3)]);
        }
    }
    __syncthreads();
    int o_end = o0 + n;
    if (o_end > kBlock) {
        twist_lds(f.v0, lane);
        twist_lds(f.v1, lane);
        for (int q = lane; q < n; q += 64) {
            const int o = o0 + q - kBlock;
            if (o >= 0) {
                const uint64_t* v = ((o & 7) < 4) ? f.v0 : f.v1;
                mt[q] = temper(v[(o >> 3) * kLanes + (o & 3)]);
            }
        }
        o_end -= kBlock;
        __syncthreads();
    }
    f.idx = (o_end + 7) >> 3;
    f.b = o_end & 7;
    ++f.refills;
    // pack the scalars (now in mt[0, n)) into draws, stored untempered
    if (f.mode == kP8F34) {
        uint64_t w[5][4];
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            const int v = lane + 64 * k;  // draw v takes scalars 3v, 3v+1, 3v+2
            if (v < kN) {
                const uint64_t a = mt[3 * v], bb = mt[3 * v + 1], c = mt[3 * v + 2];
                const uint32_t al = (uint32_t)a, ah = (uint32_t)(a >> 32), bl = (uint32_t)bb,
                               bh = (uint32_t)(bb >> 32), cl = (uint32_t)c, ch = (uint32_t)(c >> 32);
                const uint32_t x1 = (al >> 24) | ((ah >> 24) << 8) | ((bl >> 24) << 16) | ((bl >> 24) << 24);
                const uint32_t x2 = (bh >> 24) | ((cl >> 24) << 8) | ((ch >> 24) << 16) | ((ch >> 24) << 24);
                w[k][0] = a;
                w[k][1] = (uint64_t)bl | ((uint64_t)x1 << 32);
                w[k][2] = (uint64_t)bh | ((uint64_t)cl << 32);
                w[k][3] = (uint64_t)ch | ((uint64_t)x2 << 32);
            }
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            const int v = lane + 64 * k;
            if (v < kN)
#pragma unroll
                for (int l = 0; l < 4; ++l) mt[v * kLanes + l] = untemper(w[k][l]);
        }
    } else {
        // raw: draw v = scalars 4v..4v+3 in order; H2: half draw h = scalars
        // (2h, 2h+1) with lane 0 the second
        constexpr int K = (kWords + 63) / 64;
        uint64_t x[K];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int q = lane + 64 * k;
            if (q < kWords) x[k] = (f.mode == kP8H2) ? mt[q ^ 1] : mt[q];
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int q = lane + 64 * k;
            if (q < kWords) mt[q] = untemper(x[k]);
        }
    }
    __syncthreads();
}

// The next block of draws: the MT twist, or a refill from a scalar engine.
__device__ __forceinline__ void eng_refill(Eng& e, int lane) {
    if (e.feed) {
        p8_refill(*e.feed, e.mt, lane);
    } else {
        twist_lds(e.mt, lane);
    }
}

__device__ __forceinline__ void eng_load(Eng& e, uint64_t* mt, const EngineState* st, int lane) {
    load_state(mt, st, lane);
    e.mt = mt;
    e.idx = st->idx;
    e.half_valid = st->half_valid;
    e.half_cache[0] = st->half_cache[0];
    e.half_cache[1] = st->half_cache[1];
    e.twisted = false;
    e.feed = nullptr;
}

__device__ __forceinline__ void eng_store(const Eng& e, EngineState* st, int lane) {
    __syncthreads();
    if (e.twisted) store_state(st, e.mt, lane);
    if (lane == 0) {
        st->idx = e.idx;
        st->half_valid = e.half_valid;
        st->half_cache[0] = e.half_cache[0];
        st->half_cache[1] = e.half_cache[1];
    }
}

// ---------------------------------------------------------------- fixed ---
// D: CALLS draws per unit; units(per); emit(per, u, l, w) writes the outputs
// of unit u from the tempered words w[0..CALLS) of lane l. A distribution
// that reads only u64 lanes 0..LANES-1 of each draw (D::LANES = 2: the SSE2 /
// NEON layout, whose 2-lane engine is lanes 0-1 of the 4-lane record) gets
// items for those lanes only; the draws advance as before.
template <class D, class = void>
struct LanesOf {
    static constexpr int v = kLanes;
};
template <class D>
struct LanesOf<D, decltype((void)D::LANES)> {
    static constexpr int v = D::LANES;
};

template <class D>
__device__ void op_fixed(Eng& e, uint64_t* carry, int64_t per, const D& d, int lane) {
    constexpr int C = D::CALLS;
    constexpr int LU = LanesOf<D>::v;
    constexpr int MAXC = C > 1 ? C - 1 : 1;
    const int64_t units = D::units(per);
    if (units == 0) return;
    int pos = e.idx;
    int nc = 0;
    int64_t done = 0;
    for (;;) {
        if (pos >= kN) {
            eng_refill(e, lane);
            pos = 0;
            e.twisted = true;
        }
        const int avail = nc + (kN - pos);
        const int64_t can = (int64_t)(avail / C);
        const int nu = (int)(units - done < can ? units - done : can);
        int t = lane;
        if (nc > 0) {
            // Only unit 0 can hold carried words (nc < C); the first pass of
            // the wave takes the general path.
            if (t < nu * LU) {
                const int u = t / LU, l = t % LU;
                uint64_t w[C];
#pragma unroll
                for (int c = 0; c < C; ++c) w[c] = slot_word<kLanes>(e.mt, carry, nc, pos, u * C + c, l);
                d.emit(per, done + u, l, w);
            }
            t += 64;
        }
        // Units past the carried words read the block directly: word c of
        // unit u is block draw pos - nc + u*C + c.
        for (; t < nu * LU; t += 64) {
            const int u = t / LU, l = t % LU;
            uint64_t w[C];
#pragma unroll
            for (int c = 0; c < C; ++c) w[c] = temper(e.mt[(pos - nc + u * C + c) * kLanes + l]);
            d.emit(per, done + u, l, w);
        }
        const int consumed = nu * C;
        done += nu;
        if (done == units) {
            pos = pos + consumed - nc;
            break;
        }
        // Keep the slots of the unit that straddles the block boundary.
        carry_tail<kLanes, MAXC>(e.mt, carry, nc, pos, consumed, avail, lane);
        nc = avail - consumed;
        pos = kN;
    }
    __syncthreads();
    e.idx = pos;
}

// op_fixed with items of two u64 lanes: item k of unit u reads lanes 2k and
// 2k+1 of each of the unit's C draws (one 16-byte LDS read per draw), so a
// float fill writes four consecutive outputs per draw and store. Same draws,
// same order, same final engine state as op_fixed. The wave takes the units
// of a block 32 at a time (thread t: unit t/2, item t%2), with the loop and
// its bounds uniform.
// D::emit2(per, u, k, w, full): w[c][j] = tempered lane 2k+j of draw c of
// unit u; full = unit u lies wholly inside the fill (D::full_units).
template <class D>
__device__ void op_fixed2(Eng& e, uint64_t* carry, int64_t per, const D& d, int lane) {
    constexpr int C = D::CALLS;
    constexpr int MAXC = C > 1 ? C - 1 : 1;
    const int64_t units = D::units(per);
    if (units == 0) return;
    const int64_t full_units = D::full_units(per);
    int pos = e.idx;
    int nc = 0;
    int64_t done = 0;
    const int ul = lane >> 1, k = lane & 1;
    for (;;) {
        if (pos >= kN) {
            eng_refill(e, lane);
            pos = 0;
            e.twisted = true;
        }
        const int avail = nc + (kN - pos);
        const int64_t can = (int64_t)(avail / C);
        const int nu = (int)(units - done < can ? units - done : can);
        // units of this block that are whole (only the fill's last unit can be partial)
        const int nfull = (int)(full_units - done < (int64_t)nu ? full_units - done : (int64_t)nu);
        for (int cb = 0; cb < nu; cb += 32) {
            const int u = cb + ul;
            if (u < nu) {
                uint64_t w[C][2];
                if (nc > 0 && cb == 0) {
                    // only unit 0 can hold carried words
#pragma unroll
                    for (int c = 0; c < C; ++c)
#pragma unroll
                        for (int j = 0; j < 2; ++j) w[c][j] = slot_word<kLanes>(e.mt, carry, nc, pos, u * C + c, 2 * k + j);
                } else {
#pragma unroll
                    for (int c = 0; c < C; ++c) {
                        const ulonglong2 p = *reinterpret_cast<const ulonglong2*>(e.mt + (pos - nc + u * C + c) * kLanes + 2 * k);
                        w[c][0] = temper(p.x);
                        w[c][1] = temper(p.y);
                    }
                }
                d.emit2(per, done + u, k, w, u < nfull);
            }
        }
        const int consumed = nu * C;
        done += nu;
        if (done == units) {
            pos = pos + consumed - nc;
            break;
        }
        carry_tail<kLanes, MAXC>(e.mt, carry, nc, pos, consumed, avail, lane);
        nc = avail - consumed;
        pos = kN;
    }
    __syncthreads();
    e.idx = pos;
}

// op_fixed for one engine served by a workgroup of NT threads (NT/64
// waves): the units of a block are shared by all threads, wave 0 twists and
// keeps the carried words. The items and their words are those of op_fixed,
// so the outputs and the final engine state are identical.
template <int NT, class D>
__device__ void op_fixed_wg(Eng& e, uint64_t* carry, int64_t per, const D& d, int tid) {
    constexpr int C = D::CALLS;
    constexpr int MAXC = C > 1 ? C - 1 : 1;
    const int64_t units = D::units(per);
    if (units == 0) return;
    const int wl = tid < 64 ? tid : (1 << 20);  // carry_tail lane: only wave 0 moves words
    int pos = e.idx;
    int nc = 0;
    int64_t done = 0;
    for (;;) {
        if (pos >= kN) {
            twist_lds_wg<NT>(e.mt, tid);
            pos = 0;
            e.twisted = true;
        }
        const int avail = nc + (kN - pos);
        const int64_t can = (int64_t)(avail / C);
        const int nu = (int)(units - done < can ? units - done : can);
        int t = tid;
        if (nc > 0) {
            if (t < nu * kLanes) {
                const int u = t >> 2, l = t & 3;
                uint64_t w[C];
#pragma unroll
                for (int c = 0; c < C; ++c) w[c] = slot_word<kLanes>(e.mt, carry, nc, pos, u * C + c, l);
                d.emit(per, done + u, l, w);
            }
            t += NT;
        }
        for (; t < nu * kLanes; t += NT) {
            const int u = t >> 2, l = t & 3;
            uint64_t w[C];
#pragma unroll
            for (int c = 0; c < C; ++c) w[c] = temper(e.mt[(pos - nc + u * C + c) * kLanes + l]);
            d.emit(per, done + u, l, w);
        }
        const int consumed = nu * C;
        done += nu;
        if (done == units) {
            pos = pos + consumed - nc;
            break;
        }
        // carry_tail's barriers also end every wave's reads of this block
        // before the next twist
        carry_tail<kLanes, MAXC>(e.mt, carry, nc, pos, consumed, avail, wl);
        nc = avail - consumed;
        pos = kN;
    }
    __syncthreads();
    e.idx = pos;
}

// op_fixed with a per-launch draw count C (1 <= C <= MAXC + 1): D::emit gets
// fetch(c), the tempered word of draw c of its unit, instead of a word array.
template <int MAXC, class D>
__device__ void op_fixed_rt(Eng& e, uint64_t* carry, int64_t per, int C, const D& d, int lane) {
    const int64_t units = D::units(per);
    if (units == 0) return;
    int pos = e.idx;
    int nc = 0;
    int64_t done = 0;
    for (;;) {
        if (pos >= kN) {
            eng_refill(e, lane);
            pos = 0;
            e.twisted = true;
        }
        const int avail = nc + (kN - pos);
        const int64_t can = (int64_t)(avail / C);
        const int nu = (int)(units - done < can ? units - done : can);
        for (int t = lane; t < nu * kLanes; t += 64) {
            const int u = t >> 2, l = t & 3;
            d.emit(per, done + u, l, [&](int c) { return slot_word<kLanes>(e.mt, carry, nc, pos, u * C + c, l); });
        }
        const int consumed = nu * C;
        done += nu;
        if (done == units) {
            pos = pos + consumed - nc;
            break;
        }
        // Keep the draws of the unit that straddles the block boundary.
        carry_tail<kLanes, MAXC>(e.mt, carry, nc, pos, consumed, avail, lane);
        nc = avail - consumed;
        pos = kN;
    }
    __syncthreads();
    e.idx = pos;
}

// op_fixed_rt for one engine served by a workgroup of NT threads (as
// op_fixed_wg: the units of a block are shared by all threads, wave 0 twists
// and keeps the carried words); same items, outputs and final state.
template <int NT, int MAXC, class D>
__device__ void op_fixed_rt_wg(Eng& e, uint64_t* carry, int64_t per, int C, const D& d, int tid) {
    const int64_t units = D::units(per);
    if (units == 0) return;
    const int wl = tid < 64 ? tid : (1 << 20);  // carry_tail lane: only wave 0 moves words
    int pos = e.idx;
    int nc = 0;
    int64_t done = 0;
    for (;;) {
        if (pos >= kN) {
            twist_lds_wg<NT>(e.mt, tid);
            pos = 0;
            e.twisted = true;
        }
        const int avail = nc + (kN - pos);
        const int64_t can = (int64_t)(avail / C);
        const int nu = (int)(units - done < can ? units - done : can);
        for (int t = tid; t < nu * kLanes; t += NT) {
            const int u = t >> 2, l = t & 3;
            d.emit(per, done + u, l, [&](int c) { return slot_word<kLanes>(e.mt, carry, nc, pos, u * C + c, l); });
        }
        const int consumed = nu * C;
        done += nu;
        if (done == units) {
            pos = pos + consumed - nc;
            break;
        }
        // carry_tail's barriers also end every wave's reads of this block
        // before the next twist
        carry_tail<kLanes, MAXC>(e.mt, carry, nc, pos, consumed, avail, wl);
        nc = avail - consumed;
        pos = kN;
    }
    __syncthreads();
    e.idx = pos;
}

// ------------------------------------------------------------- half state ---
// Position of a half-draw consumer: block slot `pos` (in half draws) plus an
// optional carried (already tempered) slot holding the engine's cached half.
__device__ __forceinline__ void half_begin(const Eng& e, uint64_t* carry, int& pos, int& nc,
                                          int lane) {
    pos = 2 * e.idx;
    nc = 0;
    if (e.half_valid) {
        if (lane < 2) carry[lane] = e.half_cache[lane];
        nc = 1;
    }
    __syncthreads();
}

// Write back the half-draw position: an odd block slot means the draw's upper
// half is cached. When the position is even the reference still keeps the
// last split draw's upper half in its cache member; mirror that too.
__device__ __forceinline__ void half_end(Eng& e, int pos, int nc, bool used) {
    if (nc > 0) {
        e.idx = pos / 2;  // the cached half was not consumed
        return;
    }
    if (pos & 1) {
        e.idx = (pos + 1) / 2;
        e.half_valid = 1;
        e.half_cache[0] = temper(e.mt[2 * pos]);
        e.half_cache[1] = temper(e.mt[2 * pos + 1]);
    } else {
        e.idx = pos / 2;
        e.half_valid = 0;
        if (used) {
            e.half_cache[0] = temper(e.mt[2 * (pos - 1)]);
            e.half_cache[1] = temper(e.mt[2 * (pos - 1) + 1]);
        }
    }
}

// ---------------------------------------------------------------- reject ---
// The rejection cap (candidates without output before giving up) is per
// launch: Guard::reject_cap.

// Scratch a rejection op needs in LDS.
template <class P>
struct RejectLds {
    uint64_t carry[(P::K - 1) * P::SW];
    float rest[P::W];
    int32_t tab[64 / (P::W / 2) + 1];
};

// P: W lanes per packet, SW words per slot (4 full draw / 2 half draw), K
// slots per candidate; eval(w, h, v, acc) computes lane 2p+h of a candidate
// from word p of each of its K slots; store(out, i, v) writes element i.
//
// Queue order of CompressMask (arch/AVX/PacketFilter.h:147-184, arch/SSE/
// PacketFilter.h:118-150) from wave ballots: a fully accepted candidate is
// emitted as its own packet; accepted lanes of the others form the queue in
// (candidate, lane) order and queue item f lands in queued packet f/W; the
// output index of a packet = bypassing candidates before its emitter + queued
// packets before it.
// gs (optional): a reused generator's queue, taken as the initial rest queue
// and replaced by the leftovers at the end (NULL: a fresh generator whose
// leftovers are dropped, as the reference's free functions do).
template <class P>
__device__ void op_reject(Eng& e, RejectLds<P>& s, typename P::Out* o, int64_t per, const P& pol,
                          Guard gd, int lane, GenState* gs = nullptr) {
    constexpr int W = P::W, SW = P::SW, K = P::K;
    constexpr int IPC = W / 2;     // threads per candidate (two lanes each)
    constexpr int CPS = 64 / IPC;  // candidates per wave step
    constexpr int SLOTS = SlotMode<SW>::kSlots;
    constexpr int MAXC = K - 1;    // straddling slots, or the cached half
    const int64_t T = per / W;
    if (T == 0) return;
    uint64_t* carry = s.carry;
    int pos, nc = 0;
    if (SW == 2) {
        half_begin(e, carry, pos, nc, lane);
    } else {
        pos = e.idx;
    }
    bool used = false;
    int64_t full_cnt = 0, fifo = 0, since_emit = 0;
    if (gs) {
        const int r = gs->count;  // < W queued values, oldest first
        if (lane < r) s.rest[lane] = gs->v[lane];
        fifo = r;
        __syncthreads();
    }
    int end_v = -1;
    const int cl = lane / IPC, p = lane % IPC;
    for (;;) {
        if (pos >= SLOTS) {
            eng_refill(e, lane);
            pos = 0;
            e.twisted = true;
            used = true;
        }
        const int avail = nc + (SLOTS - pos);
        const int ncand = avail / K;
        for (int c0 = 0; c0 < ncand; c0 += CPS) {
            const int c = c0 + cl;
            const bool live = c < ncand;
            float v[2] = {0.f, 0.f};
            bool a[2] = {false, false};
            if (live) {
                uint64_t w[K];
#pragma unroll
                for (int k = 0; k < K; ++k) w[k] = slot_word<SW>(e.mt, carry, nc, pos, c * K + k, p);
#pragma unroll
                for (int h = 0; h < 2; ++h) pol.eval(w, h, v[h], a[h], 2 * p + h);
            }
            // Queue bookkeeping over the candidates that count (`lv`), in
            // 32-bit offsets from the step's bases: queue item fifo + x is item
            // r0 + x of queued packet m0 + (r0 + x) / W (r0 = fifo % W), and
            // packet index full_cnt + m0 + e is local packet e.
            const int cb = cl * IPC;  // the candidate's first thread
            const uint64_t cmask = (IPC == 4 ? 0xFull : 0x3ull) << cb;
            const int64_t m0 = fifo / W;
            const int r0 = (int)(fifo - m0 * W);
            const int64_t pbase = full_cnt + m0;
            const int64_t prem = T - pbase;  // local packets < prem exist
            const int lim = prem < 0x40000000 ? (int)prem : 0x40000000;
            bool full, q0, q1;
            uint64_t A0, A1, FC;
            int fstart, f0, tot, emit;
            auto plan = [&](bool lv) {
                const uint64_t B0 = __ballot(lv && a[0]), B1 = __ballot(lv && a[1]);
                full = lv && ((B0 & B1 & cmask) == cmask);
                q0 = lv && a[0] && !full;
                q1 = lv && a[1] && !full;
                A0 = __ballot(q0);
                A1 = __ballot(q1);
                FC = __ballot(full && p == 0);
                // set bits of the ballots in threads below this one; the
                // candidate's first thread has the counts before the candidate
                // (quad_perm broadcast of its lane)
                f0 = r0 + (int)mbcnt64(A1, mbcnt64(A0, 0));
                const int fbt = (int)mbcnt64(FC, 0);
                constexpr int kFirst = IPC == 4 ? 0x00 : 0xA0;  // quad_perm 0,0,0,0 / 0,0,2,2
                fstart = __builtin_amdgcn_mov_dpp(f0, kFirst, 0xF, 0xF, false);
                const int fb = __builtin_amdgcn_mov_dpp(fbt, kFirst, 0xF, 0xF, false);
                tot = __popcll(A0) + __popcll(A1);
                int acnt = (q0 ? 1 : 0) + (q1 ? 1 : 0);  // summed over the candidate's threads
                acnt += __builtin_amdgcn_mov_dpp(acnt, 0xB1, 0xF, 0xF, false);       // quad_perm 1,0,3,2
                if (IPC == 4) acnt += __builtin_amdgcn_mov_dpp(acnt, 0x4E, 0xF, 0xF, false);  // 2,3,0,1
                emit = -1;
                if (lv && (full || (acnt > 0 && (unsigned)(fstart + acnt) / W > (unsigned)fstart / W)))
                    emit = fb + (int)((unsigned)fstart / W);
            };
            plan(live);
            const uint64_t LAST = __ballot(emit == lim - 1 && p == 0);
            if (LAST) {
                // Candidates after the one emitting the last packet are never
                // drawn: drop them so that the queue ends as the reference's.
                const int clast = (int)(__ffsll((long long)LAST) - 1) / IPC;
                plan(live && cl <= clast);
            }
            const int m_end = (r0 + tot) / W;  // queued packets completed in this step
            if (!full && emit >= 0 && p == 0) s.tab[(unsigned)fstart / W] = (int32_t)(pbase + emit);
            __syncthreads();
            if (full && emit < lim) {
                const int64_t at = (pbase + emit) * W + 2 * p;
#pragma unroll
                for (int h = 0; h < 2; ++h) pol.store(o, at + h, v[h]);
            }
            if (m_end > 0 && lane < r0) {
                const int64_t ex = s.tab[0];
                if (ex < T) pol.store(o, ex * W + lane, s.rest[lane]);
            }
            float park[2] = {0.f, 0.f};
            int park_at[2] = {-1, -1};
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const bool q = h ? q1 : q0;
                const unsigned f = (unsigned)(f0 + (h ? (q0 ? 1 : 0) : 0));
                if (!q) continue;
                const int m = (int)(f / W);
                if (m < m_end) {
                    const int64_t ex = s.tab[m];
                    if (ex < T) pol.store(o, ex * W + (f % W), v[h]);
                } else {
                    park[h] = v[h];
                    park_at[h] = (int)(f % W);
                }
            }
            __syncthreads();
#pragma unroll
            for (int h = 0; h < 2; ++h)
                if (park_at[h] >= 0) s.rest[park_at[h]] = park[h];
            full_cnt += __popcll(FC);
            fifo += tot;
            if (LAST) {
                const int clast = (int)(__ffsll((long long)LAST) - 1) / IPC;
                end_v = (c0 + clast + 1) * K;
                break;
            }
            if (__popcll(FC) + m_end > 0) {
                since_emit = 0;
            } else {
                since_emit += CPS;
                if (since_emit > gd.reject_cap) break;
            }
        }
        if (end_v >= 0) break;
        if (since_emit > gd.reject_cap) {
            if (lane == 0) atomicOr(gd.status, kStatusIterCap);
            end_v = ncand * K;
            break;
        }
        carry_tail<SW, MAXC>(e.mt, carry, nc, pos, ncand * K, avail, lane);
        nc = avail - ncand * K;
        pos = SLOTS;
    }
    __syncthreads();
    if (gs) {  // leftovers: the fifo % W values of the unfinished queued packet
        const int r = (int)(fifo % W);
        if (lane < r) gs->v[lane] = s.rest[lane];
        if (lane == 0) gs->count = r;
    }
    if (end_v > nc) used = true;
    int bpos = pos, rem = nc;
    if (end_v >= nc) {
        bpos = pos + end_v - nc;
        rem = 0;
    }
    if (SW == 4) {
        e.idx = bpos;
    } else {
        half_end(e, bpos, rem, used);
    }
}

// op_reject for one engine served by NW waves (a workgroup of 64 * NW
// threads, full-draw policies, a fresh generator): a step evaluates NW * CPS
// consecutive candidates, wave w the w-th run of CPS, and each wave places
// its queue items and bypassing packets after those of the waves before it,
// so the outputs, their order and the final engine position are op_reject's.
// The serial chain of one engine's rejection fill (a Dirichlet instance's
// category rows) then runs at NW waves' rate. Wave 0 twists and carries.
template <class P, int NW>
struct RejectWgLds {
    uint64_t carry[(P::K - 1) * P::SW];
    float rest[P::W];
    int32_t tab[NW * (64 / (P::W / 2)) + 1];
    int32_t wtot[NW], wfc[NW], wlast[NW];
};

template <class P, int NW>
__device__ void op_reject_wg(Eng& e, RejectWgLds<P, NW>& s, typename P::Out* o, int64_t per, const P& pol,
                             Guard gd, int tid) {
    static_assert(P::SW == 4, "full-draw policies only");
    constexpr int W = P::W, SW = P::SW, K = P::K;
    constexpr int IPC = W / 2;         // threads per candidate (two lanes each)
    constexpr int CPS = 64 / IPC;      // candidates per wave step
    constexpr int CPSW = NW * CPS;     // candidates per workgroup step
    constexpr int SLOTS = SlotMode<SW>::kSlots;
    constexpr int MAXC = K - 1;
    const int64_t T = per / W;
    if (T == 0) return;
    const int wave = tid >> 6, lane = tid & 63;
    const int wl = tid < 64 ? tid : (1 << 20);  // carry_tail lane: only wave 0 moves words
    uint64_t* carry = s.carry;
    int pos = e.idx, nc = 0;
    int64_t full_cnt = 0, fifo = 0, since_emit = 0;
    int end_v = -1;
    const int cl = lane / IPC, p = lane % IPC;
    constexpr int kFirst = IPC == 4 ? 0x00 : 0xA0;  // quad_perm 0,0,0,0 / 0,0,2,2
    for (;;) {
        if (pos >= SLOTS) {
            twist_lds_wg<64 * NW>(e.mt, tid);
            pos = 0;
            e.twisted = true;
        }
        const int avail = nc + (SLOTS - pos);
        const int ncand = avail / K;
        for (int c0 = 0; c0 < ncand; c0 += CPSW) {
            const int c = c0 + wave * CPS + cl;
            const bool live = c < ncand;
            float v[2] = {0.f, 0.f};
            bool a[2] = {false, false};
            if (live) {
                uint64_t w[K];
#pragma unroll
                for (int k = 0; k < K; ++k) w[k] = slot_word<SW>(e.mt, carry, nc, pos, c * K + k, p);
#pragma unroll
                for (int h = 0; h < 2; ++h) pol.eval(w, h, v[h], a[h], 2 * p + h);
            }
            const int64_t m0 = fifo / W;
            const int r0 = (int)(fifo - m0 * W);
            const int64_t pbase = full_cnt + m0;
            const int64_t prem = T - pbase;  // local packets < prem exist
            const int lim = prem < 0x40000000 ? (int)prem : 0x40000000;
            const int cb = cl * IPC;
            const uint64_t cmask = (IPC == 4 ? 0xFull : 0x3ull) << cb;
            // this wave's counts, as op_reject's plan() computes them
            bool full, q0, q1;
            uint64_t FC;
            int f0l, fbl, acnt, tot_w;
            auto count = [&](bool lv) {
                const uint64_t B0 = __ballot(lv && a[0]), B1 = __ballot(lv && a[1]);
                full = lv && ((B0 & B1 & cmask) == cmask);
                q0 = lv && a[0] && !full;
                q1 = lv && a[1] && !full;
                const uint64_t A0 = __ballot(q0), A1 = __ballot(q1);
                FC = __ballot(full && p == 0);
                f0l = (int)mbcnt64(A1, mbcnt64(A0, 0));
                fbl = (int)mbcnt64(FC, 0);
                tot_w = __popcll(A0) + __popcll(A1);
                acnt = (q0 ? 1 : 0) + (q1 ? 1 : 0);  // summed over the candidate's threads
                acnt += __builtin_amdgcn_mov_dpp(acnt, 0xB1, 0xF, 0xF, false);       // quad_perm 1,0,3,2
                if (IPC == 4) acnt += __builtin_amdgcn_mov_dpp(acnt, 0x4E, 0xF, 0xF, false);  // 2,3,0,1
            };
            count(live);
            if (lane == 0) {
                s.wtot[wave] = tot_w;
                s.wfc[wave] = __popcll(FC);
            }
            __syncthreads();
            int tb = 0, fcb = 0;  // queue items and bypassing packets of the waves before this one
            for (int w2 = 0; w2 < wave; ++w2) {
                tb += s.wtot[w2];
                fcb += s.wfc[w2];
            }
            int f0, fstart, emit;
            auto place = [&](bool lv) {
                f0 = r0 + tb + f0l;
                fstart = __builtin_amdgcn_mov_dpp(f0, kFirst, 0xF, 0xF, false);
                const int fb = __builtin_amdgcn_mov_dpp(fcb + fbl, kFirst, 0xF, 0xF, false);
                emit = -1;
                if (lv && (full || (acnt > 0 && (unsigned)(fstart + acnt) / W > (unsigned)fstart / W)))
                    emit = fb + (int)((unsigned)fstart / W);
            };
            place(live);
            if (!full && emit >= 0 && emit < lim && p == 0) s.tab[(unsigned)fstart / W] = (int32_t)(pbase + emit);
            const uint64_t LAST = __ballot(emit == lim - 1 && p == 0);
            const int clast = LAST ? (int)(__ffsll((long long)LAST) - 1) / IPC : -1;
            if (lane == 0) s.wlast[wave] = clast + 1;
            __syncthreads();
            int lw = NW;  // the wave whose candidate emits the last packet (NW: none)
            for (int w2 = NW - 1; w2 >= 0; --w2)
                if (s.wlast[w2] > 0) lw = w2;
            const int lw_c = lw < NW ? s.wlast[lw] : 0;  // candidates of wave lw still drawn
            if (lw < NW) {
                // Candidates after the one emitting the last packet are never
                // drawn: drop them so that the queue ends as the reference's.
                const bool lv = live && (wave < lw || (wave == lw && cl < lw_c));
                count(lv);
                place(lv);
                if (lane == 0) {
                    s.wtot[wave] = tot_w;
                    s.wfc[wave] = __popcll(FC);
                }
                __syncthreads();
            }
            int tot = 0, nfull = 0;
            for (int w2 = 0; w2 < NW; ++w2) {
                tot += s.wtot[w2];
                nfull += s.wfc[w2];
            }
            const int m_end = (r0 + tot) / W;  // queued packets completed in this step
            if (full && emit < lim) {
                const int64_t at = (pbase + emit) * W + 2 * p;
#pragma unroll
                for (int h = 0; h < 2; ++h) pol.store(o, at + h, v[h]);
            }
            if (m_end > 0 && tid < r0) {
                const int64_t ex = s.tab[0];
                if (ex < T) pol.store(o, ex * W + tid, s.rest[tid]);
            }
            float park[2] = {0.f, 0.f};
            int park_at[2] = {-1, -1};
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const bool q = h ? q1 : q0;
                const unsigned f = (unsigned)(f0 + (h ? (q0 ? 1 : 0) : 0));
                if (!q) continue;
                const int m = (int)(f / W);
                if (m < m_end) {
                    const int64_t ex = s.tab[m];
                    if (ex < T) pol.store(o, ex * W + (f % W), v[h]);
                } else {
                    park[h] = v[h];
                    park_at[h] = (int)(f % W);
                }
            }
            __syncthreads();
#pragma unroll
            for (int h = 0; h < 2; ++h)
                if (park_at[h] >= 0) s.rest[park_at[h]] = park[h];
            full_cnt += nfull;
            fifo += tot;
            if (lw < NW) {
                end_v = (c0 + lw * CPS + lw_c) * K;
                break;
            }
            if (nfull + m_end > 0) {
                since_emit = 0;
            } else {
                since_emit += CPSW;
                if (since_emit > gd.reject_cap) break;
            }
        }
        if (end_v >= 0) break;
        if (since_emit > gd.reject_cap) {
            if (tid == 0) atomicOr(gd.status, kStatusIterCap);
            end_v = ncand * K;
            break;
        }
        carry_tail<SW, MAXC>(e.mt, carry, nc, pos, ncand * K, avail, wl);
        nc = avail - ncand * K;
        pos = SLOTS;
    }
    __syncthreads();
    e.idx = end_v >= nc ? pos + end_v - nc : pos;
}

// -------------------------------------------------------------- lockstep ---
constexpr int kHalfSlots = kWords / 2;  // half draws per block
// The lockstep cap (iterations of one packet before giving up) is per launch:
// Guard::lockstep_cap.

// Cursor over the slot sequence of a lockstep op: SW words per slot (4 full
// draw, 2 half draw), at most MAXC carried slots across a twist.
template <int SW, int MAXC>
struct Cursor {
    static constexpr int kSlots = SlotMode<SW>::kSlots;
    uint64_t* mt;
    uint64_t* carry;
    int pos;
    int nc;
    bool used;
    Eng* eng;  // its refill and twisted flag

    __device__ int avail() const { return nc + (kSlots - pos); }
    __device__ uint64_t word(int v, int p) const { return slot_word<SW>(mt, carry, nc, pos, v, p); }
    // Tempered 32-bit word of float lane j (0..2*SW-1) in virtual slot v.
    __device__ uint32_t word32(int v, int j) const {
        const uint64_t w = word(v, j >> 1);
        return (j & 1) ? hi32(w) : lo32(w);
    }
    __device__ void advance(int n) {
        if (n >= nc) {
            if (n > nc) used = true;
            pos += n - nc;
            nc = 0;
        } else {
            // Only part of the carried slots consumed: shift them down.
            for (int i = 0; i + n < nc; ++i)
                for (int q = 0; q < SW; ++q) carry[i * SW + q] = carry[(i + n) * SW + q];
            nc -= n;
        }
    }
    // Make at least `need` (<= MAXC+1) slots available; called by the whole wave.
    __device__ void ensure(int need, int lane) {
        const int a = avail();
        if (a >= need) return;
        carry_tail<SW, MAXC>(mt, carry, nc, pos, 0, a, lane);
        nc = a;
        eng_refill(*eng, lane);
        pos = 0;
        eng->twisted = true;
    }
};
using HalfCursor = Cursor<2, 1>;

__device__ __forceinline__ HalfCursor cursor_begin(Eng& e, uint64_t* carry, int lane) {
    HalfCursor c;
    c.mt = e.mt;
    c.carry = carry;
    c.used = false;
    c.eng = &e;
    half_begin(e, carry, c.pos, c.nc, lane);
    return c;
}

__device__ __forceinline__ void cursor_end(Eng& e, const HalfCursor& c) {
    __syncthreads();
    half_end(e, c.pos, c.nc, c.used);
}

// Fixed half-draw consumers (D::SLOTS half() draws per Packet4i): packet k
// of the fill takes virtual slots [k*SLOTS, (k+1)*SLOTS) (the engine's cached
// half first, if any), 64 packets per wave step; a last odd half leaves the
// draw's upper half cached. emit(i, x) gets the tempered words x[2*s + q]
// (slot s, word q).
template <class D>
__device__ void op_half_fixed(Eng& e, uint64_t* carry, int64_t T, const D& d, int lane) {
    constexpr int S = D::SLOTS;
    if (T == 0) return;
    HalfCursor cur = cursor_begin(e, carry, lane);
    int64_t done = 0;
    while (done < T) {
        cur.ensure(S, lane);
        const int64_t take = min((int64_t)(cur.avail() / S), T - done);
        for (int v = lane; v < take; v += 64) {
            uint64_t x[2 * S];
#pragma unroll
            for (int k = 0; k < S; ++k) {
                x[2 * k] = cur.word(v * S + k, 0);
                x[2 * k + 1] = cur.word(v * S + k, 1);
            }
            d.emit(4 * (done + v), x);
        }
        __syncthreads();  // carried slots are read before ensure() rewrites them
        cur.advance((int)take * S);
        done += take;
    }
    cursor_end(e, cur);
}

// UniformIntGen<int32_t>::packetOp, rejection branch (Discrete.h:291-335):
// a packet draws a half, tests its candidates (rx & mask, then rx shifted
// right by `bits` while bits fit) through CompressMask<16> (W = 4) and, once a
// packet is emitted, drops the rest of that draw; the next packet starts on a
// new half. So every half draw ("slot") emits at most one packet, and the
// process is a chain over slots whose state is the queue count r (0..3).
//
// A wave takes 64 slots at once: thread t evaluates its slot's candidates and
// the slot's transition for each incoming r (emission, candidate, values
// appended, outgoing r); a prefix composition of the transitions over the
// wave gives every slot its actual incoming r. Queued values live in an LDS
// ring at consecutive positions, so the queue at any point is the last r
// values appended and a queue emission outputs the 4 values starting at the
// oldest queued one. The fill stops at the slot emitting its last packet.
struct UniformIntParams {
    uint32_t mask;  // bitmask as 32 bits
    int32_t plen;   // pdiff + 1 as int32 (compared signed, as pcmplt does)
    int32_t pmin;
    int bits;       // bitsize
    int cands;      // candidates per half draw
};
constexpr int kUiRing = 512;  // >= 64 slots x 6 appended values + 3 queued

__device__ void op_uniform_int(Eng& e, uint64_t* carry, int32_t* ring, int32_t* o, int64_t T,
                               const UniformIntParams& u, int lane) {
    if (T == 0) return;
    HalfCursor cur = cursor_begin(e, carry, lane);
    int64_t pk = 0;
    int r = 0;          // queued values (wave-uniform)
    uint32_t head = 0;  // ring position after the newest queued value
    while (pk < T) {
        cur.ensure(1, lane);
        const int take = min(cur.avail(), 64);
        const bool live = lane < take;
        uint32_t w[4] = {0u, 0u, 0u, 0u};
        if (live) {
            const uint64_t x0 = cur.word(lane, 0), x1 = cur.word(lane, 1);
            w[0] = lo32(x0); w[1] = hi32(x0); w[2] = lo32(x1); w[3] = hi32(x1);
        }
        uint64_t accm = 0;  // acceptance bits: 4 per candidate
        for (int k = 0; k < u.cands; ++k)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int32_t c = (int32_t)((w[j] >> (k * u.bits)) & u.mask);
                if (c < u.plen) accm |= 1ull << (4 * k + j);
            }
        // transition per incoming count: out count (2 bits), kind (0 none,
        // 1 bypass, 2 queue), emitting candidate, values appended
        uint32_t fmap = 0, kinds = 0, kem = 0, apps = 0;
#pragma unroll
        for (int r0 = 0; r0 < 4; ++r0) {
            int rr = r0, a = 0, kind = 0, ke = 0;
            if (live) {
                for (int k = 0; k < u.cands; ++k) {
                    const uint32_t m = (uint32_t)(accm >> (4 * k)) & 0xFu;
                    if (m == 0xFu) { kind = 1; ke = k; break; }
                    const int n = __popc(m);
                    a += n;
                    rr += n;
                    if (rr >= 4) { kind = 2; ke = k; rr -= 4; break; }
                }
            }
            fmap |= (uint32_t)rr << (2 * r0);
            kinds |= (uint32_t)kind << (2 * r0);
            kem |= (uint32_t)ke << (5 * r0);
            apps |= (uint32_t)a << (3 * r0);
        }
        // inclusive prefix composition of the maps over the wave (slot order)
        uint32_t pre = fmap;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t before = (uint32_t)__shfl_up((int)pre, d);
            if (lane >= d) {
                uint32_t h = 0;
#pragma unroll
                for (int q = 0; q < 4; ++q) h |= ((pre >> (2 * ((before >> (2 * q)) & 3u))) & 3u) << (2 * q);
                pre = h;
            }
        }
        uint32_t excl = (uint32_t)__shfl_up((int)pre, 1);
        if (lane == 0) excl = 0xE4u;  // identity: 0,1,2,3
        const int rin = (int)((excl >> (2 * r)) & 3u);
        const int kind = (int)((kinds >> (2 * rin)) & 3u);
        const int ke = (int)((kem >> (5 * rin)) & 31u);
        const int a = (int)((apps >> (3 * rin)) & 7u);
        const int rout = (int)((fmap >> (2 * rin)) & 3u);
        // packets: slot t emits packet pk + (emitting slots before t); the
        // fill ends at the slot that emits packet T - 1
        const uint64_t E = __ballot(kind != 0);
        const int64_t before = __popcll(E & lanemask_lt(lane));
        const int64_t need = T - pk;
        int used = take;
        if ((int64_t)__popcll(E) >= need) {
            uint64_t m = E;
            for (int64_t q = 1; q < need; ++q) m &= m - 1;
            used = (int)__ffsll((long long)m);  // slots up to the emitting one
        }
        const bool mine = lane < used;
        // ring positions of appended values: exclusive prefix over the wave
        int incl = mine ? a : 0;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int v = __shfl_up(incl, d);
            if (lane >= d) incl += v;
        }
        const uint32_t A = head + (uint32_t)(incl - (mine ? a : 0));
        if (mine) {
            uint32_t pos = A;
            const int last = kind == 1 ? ke - 1 : (kind == 2 ? ke : u.cands - 1);
            for (int k = 0; k <= last; ++k) {
                const uint32_t m = (uint32_t)(accm >> (4 * k)) & 0xFu;
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    if (m & (1u << j)) ring[pos++ % kUiRing] = (int32_t)((w[j] >> (k * u.bits)) & u.mask);
            }
        }
        __syncthreads();
        if (mine && kind != 0) {
            int4 v;
            if (kind == 1) {
                v = make_int4((int32_t)((w[0] >> (ke * u.bits)) & u.mask), (int32_t)((w[1] >> (ke * u.bits)) & u.mask),
                              (int32_t)((w[2] >> (ke * u.bits)) & u.mask), (int32_t)((w[3] >> (ke * u.bits)) & u.mask));
            } else {
                const uint32_t q = A - (uint32_t)rin;
                v = make_int4(ring[q % kUiRing], ring[(q + 1) % kUiRing], ring[(q + 2) % kUiRing],
                              ring[(q + 3) % kUiRing]);
            }
            v.x += u.pmin; v.y += u.pmin; v.z += u.pmin; v.w += u.pmin;
            *reinterpret_cast<int4*>(o + 4 * (pk + before)) = v;
        }
        r = __builtin_amdgcn_readlane(rout, used - 1);
        head += (uint32_t)__builtin_amdgcn_readlane(incl, used - 1);
        pk += __popcll(E & (used == 64 ? ~0ull : ((1ull << used) - 1)));
        __syncthreads();  // ring and carried slots are rewritten next
        cur.advance(used);
    }
    cursor_end(e, cur);
}

// Knuth loop of PoissonGen::packetOp (Discrete.h:860-873); returns lane j's count.
__device__ __forceinline__ int32_t knuth_packet(HalfCursor& cur, float thr, int lane,
                                                Guard gd) {
    const int j = lane & 3;
    float val = 1.0f;
    int32_t res = 0;
    for (int it = 0;; ++it) {
        cur.ensure(1, lane);
        const float u = ur_f32(cur.word32(0, j));
        cur.advance(1);
        val = val * u;
        const bool c = thr < val;
        if (!(__ballot(c) & 0xFull)) break;
        res += c ? 1 : 0;
        if (it > gd.lockstep_cap) {
            if (lane == 0) atomicOr(gd.status, kStatusIterCap);
            break;
        }
    }
    return res;
}

// op_knuth's wide path: every thread evaluates the packet that would start at
// one of 64 consecutive slots, for up to 32 iterations, so a window stages the
// uniforms of 95 slots in LDS (16-byte aligned, one float4 per slot).
constexpr int kKnuthStarts = 64;
constexpr int kKnuthIters = 32;
constexpr int kKnuthSlots = kKnuthStarts + kKnuthIters - 1;
constexpr int kKnuthStage = kKnuthSlots * 4;

// op_knuth's prefix path (larger means, where a packet spans many slots):
// a window of kLpSlots slots holds, per lane, prefix sums of a fixed-point
// log2 of the uniforms, so the stopping iteration of a packet starting at any
// slot is one comparison per candidate iteration (one per thread).
//
// Exactness. The loop's product p_i = fl(p_{i-1} u_i) satisfies
// |log2 p_i - sum log2 u| <= (i+1) 2^-24 / ln 2 (one rounding per product;
// every product compared is a normal number above e^-12 or is the first one
// at or below the threshold). The prefix difference D_i approximates
// sum log2 u with at most (i+1) (e_log + 2^-20) error: e_log is the error of
// v_log_f32 on the values k 2^-23 (at most 1.9e-6, one ulp, measured over all
// 2^23 of them) and 2^-20 the rounding to kLpFrac fraction bits; the
// threshold's own log2 adds the same once. For i < 64 the total stays under
// 2^-12 = 128 fixed-point units, so an iteration whose D_i lies more than
// kLpMargin = 256 units below
// (above) log2(thr) has p_i <= thr (p_i > thr). Only if an iteration before
// the first certainly-stopped one lies within the margin is the packet's lane
// recomputed with the exact product loop (about 1 lane in 1500).
constexpr int kLpSlots = 128;
constexpr int kLpFrac = 19;
constexpr int32_t kLpMargin = 256;
// Thresholds e^-mean below this (mean above 4, where a packet spans about
// nine slots or more and the wide path's 64 speculative starts are mostly
// wasted; measured crossover of the two paths) take the prefix path.
constexpr float kKnuthPrefixThr = 0.0183156389f;  // e^-4

// Tempered word q of window slot v: the carried half (v < nc), block slot
// pos + v - nc, or past the block end a slot of the next block, computed
// read-only from the untwisted block (its first kM words depend on this
// block only); the block is twisted in place once a packet uses such a slot.
__device__ __forceinline__ uint64_t lp_word(const uint64_t* mt, const uint64_t* carry, int pos, int nc, int v,
                                            int q) {
    if (v < nc) return carry[q];
    const int b = pos + v - nc;
    if (b < kHalfSlots) return temper(mt[2 * b + q]);
    const int w = 2 * (b - kHalfSlots) + q;
    return temper(twist_word(mt[w], mt[w + kLanes], mt[w + kM * kLanes]));
}

__device__ __forceinline__ float lp_u(const uint64_t* mt, const uint64_t* carry, int pos, int nc, int v, int j) {
    const uint64_t x = lp_word(mt, carry, pos, nc, v, j >> 1);
    return ur_f32((j & 1) ? hi32(x) : lo32(x));
}

// Fixed-point log2 (kLpFrac fraction bits); u = 0 maps to -32, below every
// threshold's log2 (the threshold is at least e^-12).
__device__ __forceinline__ uint32_t lp_fix(float u) {
    const float l = fmaxf(__builtin_amdgcn_logf(u), -32.0f);
    return (uint32_t)__float2int_rn(l * (float)(1 << kLpFrac));
}

// Inclusive scan over the wave, modulo 2^32: Hillis-Steele within rows of 16
// lanes (DPP row shifts, zero past the row start), then the row totals.
template <int N>
__device__ __forceinline__ uint32_t dpp_row_shr(uint32_t x) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x110 + N, 0xf, 0xf, true);
}
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x, int lane) {
    x += dpp_row_shr<1>(x);
    x += dpp_row_shr<2>(x);
    x += dpp_row_shr<4>(x);
    x += dpp_row_shr<8>(x);
    const uint32_t r0 = __builtin_amdgcn_readlane(x, 15);
    const uint32_t r1 = r0 + __builtin_amdgcn_readlane(x, 31);
    const uint32_t r2 = r1 + __builtin_amdgcn_readlane(x, 47);
    return x + (lane < 16 ? 0u : lane < 32 ? r0 : lane < 48 ? r1 : r2);
}

// The window's prefix sums stay in registers: thread t holds
// r0 = P[t + 1] and r1 = P[t + 65] per lane, where P[k] is the sum of lp_fix
// over window slots < k (P[0] = 0). Sums wrap modulo 2^32; differences over
// at most 64 slots are exact.
struct LpWindow {
    uint32_t r0[4], r1[4];
};

__device__ __forceinline__ void lp_stage(const uint64_t* mt, const uint64_t* carry, int pos, int nc, LpWindow& w,
                                         int lane) {
    uint32_t a0[4], a1[4];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const uint64_t x0 = lp_word(mt, carry, pos, nc, lane, q);
        const uint64_t x1 = lp_word(mt, carry, pos, nc, lane + 64, q);
        a0[2 * q] = lp_fix(ur_f32(lo32(x0)));
        a0[2 * q + 1] = lp_fix(ur_f32(hi32(x0)));
        a1[2 * q] = lp_fix(ur_f32(lo32(x1)));
        a1[2 * q + 1] = lp_fix(ur_f32(hi32(x1)));
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        w.r0[j] = wave_incl_scan(a0[j], lane);
        w.r1[j] = __builtin_amdgcn_readlane(w.r0[j], 63) + wave_incl_scan(a1[j], lane);
    }
}

// P[s + 1 + lane] for window slot s (0 <= s <= 128): the value lives in lane
// (s + lane) mod 64, in r0 if that lane is at or after s, else in r1.
__device__ __forceinline__ uint32_t lp_fetch(const LpWindow& w, int j, int s, int lane) {
    const uint32_t v = lane < s ? w.r1[j] : w.r0[j];
    return (uint32_t)__builtin_amdgcn_ds_bpermute(((s + lane) & 63) << 2, (int)v);
}

// Exact stopping iteration of lane j for a packet starting at window slot s:
// the reference's product loop, at most `cap` iterations (wave-uniform).
__device__ __attribute__((noinline)) int lp_exact(const uint64_t* mt, const uint64_t* carry, int pos, int nc,
                                                   int s, int j, float thr, int cap) {
    float val = 1.0f;
    int k = 0;
    for (; k < cap; ++k) {
        val = val * lp_u(mt, carry, pos, nc, s + k, j);
        if (!(thr < val)) break;
    }
    return __builtin_amdgcn_readfirstlane(k);
}

// Move the cursor past `n` window slots; a position past the block end means
// a packet used slots of the next block, so the block is twisted now.
__device__ __forceinline__ void lp_advance(HalfCursor& cur, int n, int lane) {
    cur.advance(n);
    if (cur.pos > kHalfSlots) {
        __syncthreads();
        eng_refill(*cur.eng, lane);
        cur.pos -= kHalfSlots;
        cur.eng->twisted = true;
    }
}

// First index (0..63) at which lane j of the packet is certainly stopped, and
// whether an index up to it lies within the margin (then the exact loop decides).
struct LpLane {
    uint64_t stop, amb;
};
__device__ __forceinline__ LpLane lp_lane(uint32_t pt, uint32_t c, bool in) {
    // Past the window: a value that is neither stopped nor within the margin.
    const uint32_t d = in ? pt - c : 0x7fffffffu;
    LpLane r;
    r.stop = __builtin_amdgcn_ballot_w64((int32_t)d < -kLpMargin);
    r.amb = __builtin_amdgcn_ballot_w64(d + (uint32_t)kLpMargin <= (uint32_t)(2 * kLpMargin));
    return r;
}

// Prefix path of op_knuth: per window, one step per packet. Thread t tests
// iteration t of the packet starting at window slot s for all 4 lanes
// against the threshold (one ballot per lane gives each lane's first
// certainly-stopped iteration); the packet consumes up to its slowest lane
// and the next starts after it. A packet whose stop lies past the window
// restarts the window at its first slot.
__device__ void op_knuth_lp(Eng& e, uint64_t* carry, int32_t* o, int64_t T, float thr, int32_t trials, int flip,
                            Guard gd, int lane) {
    HalfCursor cur = cursor_begin(e, carry, lane);
    const uint32_t tfix = (uint32_t)__builtin_amdgcn_readfirstlane(
        __float2int_rn(__builtin_amdgcn_logf(thr) * (float)(1 << kLpFrac)));
    const int j = lane & 3;
    int64_t flushed = 0;  // packets written to o
    int nb = 0;           // packets held in obuf (lane 4k+j: packet flushed+k, lane j)
    int32_t obuf = 0;
    LpWindow w;
    while (flushed + nb < T) {
        lp_stage(cur.mt, carry, cur.pos, cur.nc, w, lane);
        int s = 0;
        uint32_t c0 = tfix, c1 = tfix, c2 = tfix, c3 = tfix;  // P[s] + log2(thr), per lane
        bool serial = false;
        while (flushed + nb < T) {
            const int room = kLpSlots - s;
            const bool in = lane < room;
            const uint32_t p0 = lp_fetch(w, 0, s, lane), p1 = lp_fetch(w, 1, s, lane);
            const uint32_t p2 = lp_fetch(w, 2, s, lane), p3 = lp_fetch(w, 3, s, lane);
            const LpLane l0 = lp_lane(p0, c0, in), l1 = lp_lane(p1, c1, in);
            const LpLane l2 = lp_lane(p2, c2, in), l3 = lp_lane(p3, c3, in);
            if (!l0.stop || !l1.stop || !l2.stop || !l3.stop) {
                serial = room >= 64;  // longer than 64 iterations: exact one-slot loop
                break;
            }
            int st0 = (int)__builtin_ctzll(l0.stop), st1 = (int)__builtin_ctzll(l1.stop);
            int st2 = (int)__builtin_ctzll(l2.stop), st3 = (int)__builtin_ctzll(l3.stop);
            if (l0.amb & ((2ull << st0) - 1))
                st0 = __builtin_amdgcn_readfirstlane(lp_exact(cur.mt, carry, cur.pos, cur.nc, s, 0, thr, st0));
            if (l1.amb & ((2ull << st1) - 1))
                st1 = __builtin_amdgcn_readfirstlane(lp_exact(cur.mt, carry, cur.pos, cur.nc, s, 1, thr, st1));
            if (l2.amb & ((2ull << st2) - 1))
                st2 = __builtin_amdgcn_readfirstlane(lp_exact(cur.mt, carry, cur.pos, cur.nc, s, 2, thr, st2));
            if (l3.amb & ((2ull << st3) - 1))
                st3 = __builtin_amdgcn_readfirstlane(lp_exact(cur.mt, carry, cur.pos, cur.nc, s, 3, thr, st3));
            const int last = max(max(st0, st1), max(st2, st3));  // consumption - 1
            const int mine = j == 0 ? st0 : j == 1 ? st1 : j == 2 ? st2 : st3;
            obuf = (lane >> 2) == nb ? mine : obuf;
            if (++nb == 16) {
                o[4 * flushed + lane] = flip ? trials - obuf : obuf;
                flushed += 16;
                nb = 0;
            }
            c0 = __builtin_amdgcn_readlane(p0, last) + tfix;
            c1 = __builtin_amdgcn_readlane(p1, last) + tfix;
            c2 = __builtin_amdgcn_readlane(p2, last) + tfix;
            c3 = __builtin_amdgcn_readlane(p3, last) + tfix;
            s += last + 1;
        }
        lp_advance(cur, s, lane);
        if (serial) {
            if (lane < 4 * nb) o[4 * flushed + lane] = flip ? trials - obuf : obuf;
            flushed += nb;
            nb = 0;
            const int32_t r = knuth_packet(cur, thr, lane, gd);
            if (lane < 4) o[4 * flushed + lane] = flip ? trials - r : r;
            ++flushed;
        }
    }
    if (lane < 4 * nb) o[4 * flushed + lane] = flip ? trials - obuf : obuf;
    cursor_end(e, cur);
}

// PoissonGen, mean < 12; BinomialGen's Poisson regime with the final flip.
//
// A packet starting at slot P consumes 1 + max_j(stop_j) half draws, where
// stop_j is the first iteration whose running product of lane j is <= thr;
// the product never increases, so lane j's count is stop_j.
//
// Wide path: the wave stages the uniforms of slots P..P+94 in LDS (or up to
// the end of the block), and thread t multiplies the 4 lanes' uniforms from
// slot P+t on, in the loop's order, for up to 32 iterations inside the
// window. That is the packet that would start at P+t. The
// packets actually produced form a chain: start 0, then each next start is
// the previous start plus its consumption. A scalar walk marks the chain
// while starts lie in the window and each packet completes within its 32
// iterations, and the marked threads store their packets at their rank in
// the chain. A packet needing more iterations falls back to the
// one-slot-at-a-time loop. Results and consumption are those of the
// sequential loop. Larger means (thr < kKnuthPrefixThr), where most of the 64
// speculative starts are wasted, take op_knuth_lp.
__device__ void op_knuth(Eng& e, uint64_t* carry, float* ustage, int32_t* o, int64_t per, float thr,
                         int32_t trials, int flip, Guard gd, int lane) {
    const int64_t T = per / 4;
    if (T == 0) return;
    if (thr < kKnuthPrefixThr) {
        op_knuth_lp(e, carry, o, T, thr, trials, flip, gd, lane);
        return;
    }
    HalfCursor cur = cursor_begin(e, carry, lane);
    int64_t pk = 0;
    while (pk < T) {
        bool slow = true;
        if (cur.nc == 0 && cur.avail() > 0) {
            // The window is the next S block slots (fewer at the end of the
            // block: packets that would need slots past it are finished by
            // the one-slot loop, which twists at the same point the
            // reference does). u64 word w holds floats 2w, 2w+1 (slot w/2).
            const int S = min(cur.avail(), kKnuthSlots);
            for (int w = lane; w < 2 * S; w += 64) {
                const uint64_t x = temper(cur.mt[2 * cur.pos + w]);
                *reinterpret_cast<float2*>(ustage + 2 * w) = make_float2(ur_f32(lo32(x)), ur_f32(hi32(x)));
            }
            __syncthreads();
            // Iterations thread `lane` may take without leaving the window.
            const int iters = max(0, min(kKnuthIters, S - lane));
            float v0 = 1.0f, v1 = 1.0f, v2 = 1.0f, v3 = 1.0f;
            int s0 = kKnuthIters, s1 = kKnuthIters, s2 = kKnuthIters, s3 = kKnuthIters;
            for (int s = 0; s < iters; ++s) {
                const float4 u = *reinterpret_cast<const float4*>(ustage + 4 * (lane + s));
                v0 = v0 * u.x;
                v1 = v1 * u.y;
                v2 = v2 * u.z;
                v3 = v3 * u.w;
                if (s0 == kKnuthIters && !(thr < v0)) s0 = s;
                if (s1 == kKnuthIters && !(thr < v1)) s1 = s;
                if (s2 == kKnuthIters && !(thr < v2)) s2 = s;
                if (s3 == kKnuthIters && !(thr < v3)) s3 = s;
                if (s0 < kKnuthIters && s1 < kKnuthIters && s2 < kKnuthIters && s3 < kKnuthIters) break;
            }
            const int m = max(max(s0, s1), max(s2, s3));
            const int cons = m < iters ? m + 1 : 0;  // 0: runs past the window
            // The chain from start 0, by pointer jumping instead of a walk.
            // Position v < L (L = starts in this window) is a start whose next
            // start is v + cons[v]; positions >= L end the chain, and so does a
            // start whose packet runs past the window (marked kOvf: the walk
            // stops before it and the one-slot loop finishes that packet).
            // Thread h ends with pos = next^h(0), the start of chain packet h
            // (next^a and next^b commute, so the bits of h apply in any order).
            constexpr int kOvf = 0x100;
            const int L = min(kKnuthStarts, S);
            auto lookup = [&](int table, int v) {  // table[v] for a start v; v itself otherwise
                const int t = __builtin_amdgcn_ds_bpermute(4 * (v & 63), table);
                return v < L ? t : v;
            };
            int jmp = lane < L ? (cons ? lane + cons : (lane | kOvf)) : lane;  // next^(2^k)[lane]
            int pos = 0;
#pragma unroll
            for (int k = 0; k < 6; ++k) {
                const int moved = lookup(jmp, pos);
                if ((lane >> k) & 1) pos = moved;
                jmp = lookup(jmp, jmp);
            }
            // a chain packet is a start whose own packet fits (a start with
            // cons 0 is where the walk stops); the live ones form a prefix
            const int cpos = __builtin_amdgcn_ds_bpermute(4 * (pos & 63), cons);
            const int64_t room = T - pk;
            const int nch = __popcll(__ballot(pos < L && cpos > 0));
            const int n = (int)(room < nch ? room : nch);
            int at = 0;  // slots consumed: just past chain packet n-1
            if (n > 0) {
                const int last = __builtin_amdgcn_readlane(pos, n - 1);
                at = last + __builtin_amdgcn_readlane(cons, last);
            }
            // the walk meets a start whose packet runs past the window
            slow = n < room && at < L && __builtin_amdgcn_readlane(cons, at) == 0;
            // chain packet h takes the counts of the thread at its start
            const int packed = s0 | (s1 << 8) | (s2 << 16) | (s3 << 24);
            const int got = __builtin_amdgcn_ds_bpermute(4 * (pos & 63), packed);
            if (lane < n) {
                const int g0 = got & 0xFF, g1 = (got >> 8) & 0xFF, g2 = (got >> 16) & 0xFF, g3 = (got >> 24) & 0xFF;
                *reinterpret_cast<int4*>(o + 4 * (pk + lane)) =
                    flip ? make_int4(trials - g0, trials - g1, trials - g2, trials - g3) : make_int4(g0, g1, g2, g3);
            }
            pk += n;
            __syncthreads();  // ustage is rewritten by the next window
            cur.advance(at);
        }
        if (slow && pk < T) {
            const int32_t r = knuth_packet(cur, thr, lane, gd);
            if (lane < 4) o[4 * pk + lane] = flip ? trials - r : r;
            ++pk;
        }
    }
    cursor_end(e, cur);
}

// BinomialGen, trials < 25 (Discrete.h:996-1006): packet k counts, per lane,
// the half draws [k*n, k*n+n) below p, so packets are independent: each thread
// takes a packet's two lanes of one u64 word. A packet straddling a twist
// keeps up to 23 slots in the carry area.
constexpr int kBinSmallMaxCarry = 23;
__device__ void op_binomial_small(Eng& e, uint64_t* carry, int32_t* o, int64_t per, int32_t trials,
                                  float pp, int lane) {
    const int64_t T = per / 4;
    const int S = trials;  // half draws per packet, < 25
    if (T == 0) return;
    if (S == 0) {  // no draws: every count is 0
        for (int64_t i = lane; i < 4 * T; i += 64) o[i] = 0;
        return;
    }
    Cursor<2, kBinSmallMaxCarry> cur;
    cur.mt = e.mt;
    cur.carry = carry;
    cur.used = false;
    cur.eng = &e;
    half_begin(e, carry, cur.pos, cur.nc, lane);
    int64_t done = 0;
    while (done < T) {
        cur.ensure(S, lane);
        const int64_t can = cur.avail() / S;
        const int take = (int)(T - done < can ? T - done : can);
        // packet q takes slots [q*S, q*S+S); thread v counts lanes 2l, 2l+1
        for (int v = lane; v < 2 * take; v += 64) {
            const int q = v >> 1, l = v & 1;
            int32_t r0 = S, r1 = S;
            for (int k = 0; k < S; ++k) {
                const uint64_t w = cur.word(q * S + k, l);
                r0 -= (pp <= ur_f32(lo32(w))) ? 1 : 0;
                r1 -= (pp <= ur_f32(hi32(w))) ? 1 : 0;
            }
            *reinterpret_cast<int2*>(o + 4 * (done + q) + 2 * l) = make_int2(r0, r1);
        }
        __syncthreads();  // carried slots are read before ensure() rewrites them
        cur.advance(take * S);
        done += take;
    }
    __syncthreads();
    half_end(e, cur.pos, cur.nc, cur.used);
}

// BinomialVGen<int32_t>::packetOp (Discrete.h:1124-1192). Tangent iterations
// are evaluated 16 at a time (thread 4i+j runs iteration i of lane j); each
// open lane takes its first accepted candidate and the packet consumes the
// iterations up to the slowest lane, as the sequential loop does.
__device__ void op_binomial_vop(Eng& e, uint64_t* carry, int32_t* o, int64_t per,
                                Param<int32_t> trials, Param<float> prob, Guard gd,
                                const float* lg, int lane) {
    const int64_t T = per / 4;
    if (T == 0) return;
    const int i = lane >> 2, j = lane & 3;
    HalfCursor cur = cursor_begin(e, carry, lane);
    const uint64_t jmask = 0x1111111111111111ull << j;
    int32_t g_n = 0;
    float g_p = 0.f, g_pg1 = 0.f, g_psq = 0.f, g_plp = 0.f, g_plq = 0.f, g_nem = 0.f;
    for (int64_t pk = 0; pk < T; ++pk) {
        if ((pk & 15) == 0) {
            // constants of packets pk..pk+15: thread t holds packet pk + t/4, lane t%4
            const bool in = pk + i < T;
            g_n = in ? trials[4 * pk + lane] : 1;
            g_p = in ? prob[4 * pk + lane] : 0.5f;
            const float gs = (g_p < 1.0f - g_p) ? g_p : 1.0f - g_p;
            const float gtr = (float)g_n;
            const float gmean = gtr * gs;
            g_pg1 = plgamma_approx(gtr + 1.0f);
            g_psq = sqrtf((2.0f * gmean) * (1.0f - gs));
            g_plp = plog(gs);
            g_plq = plog(1.0f - gs);
            g_nem = pexp(-gmean);
        }
        const int src = 4 * (int)(pk & 15) + j;
        const int32_t n = __shfl(g_n, src);
        const float p = __shfl(g_p, src);
        const float psmall = (p < 1.0f - p) ? p : 1.0f - p;  // pmin(p, 1-p)
        const float ptr = (float)n;
        const float pmean = ptr * psmall;
        const float pg1 = __shfl(g_pg1, src), psq = __shfl(g_psq, src), plp = __shfl(g_plp, src),
                    plq = __shfl(g_plq, src);
        const bool valid = pmean < 1.0f;
        float valid_res = 0.0f;
        const uint64_t V = __ballot(valid) & 0xFull;
        if (V) {
            const float nem = __shfl(g_nem, src);
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
                if (it > gd.lockstep_cap) {
                    if (lane == 0) atomicOr(gd.status, kStatusIterCap);
                    break;
                }
            }
            valid_res = (float)res;
        }
        if (V != 0xFull) {
            uint64_t done = V;
            for (int it = 0;; it += 16) {
                cur.ensure(2, lane);
                const int ni_avail = cur.avail() / 2;
                const int ni = ni_avail < 16 ? ni_avail : 16;
                float fres = 0.0f;
                bool fv = false;
                if (i < ni) {
                    float s, c;
                    psincos_pos(ERD_F(kPi) * ur_f32(cur.word32(2 * i, j)), s, c);
                    const float ys = s / c;
                    fres = truncf(psq * ys + pmean);
                    const float pp1 = (ERD_F(1.2) * psq) * (1.0f + ys * ys);
                    const bool c1 = (0.0f <= fres) && (fres <= ptr);
                    float lg1 = 0.0f, lg2 = 0.0f;  // unused when the range test fails
                    if (c1) {
                        lg1 = plgamma_int(lg, fres + 1.0f);
                        lg2 = plgamma_int(lg, (ptr + 1.0f) - fres);
                    }
                    const float pp2 = pexp((((pg1 - lg1) - lg2) + fres * plp) + (ptr - fres) * plq);
                    const float u2 = ur_f32(cur.word32(2 * i + 1, j));
                    fv = c1 && (u2 <= pp1 * pp2);
                }
                const uint64_t A = __ballot(fv);
                int need = 0;
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
                // lane j's first accepted value, from the thread that drew it
                const uint64_t mine = A & jmask;
                const float got = __shfl(fres, mine ? (int)(__ffsll((long long)mine) - 1) : lane);
                if (!(done & (1ull << j)) && mine) valid_res = got;
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    if (A & (0x1111111111111111ull << q)) done |= 1ull << q;
                cur.advance(2 * need);
                if (all) break;
                if (it > gd.lockstep_cap) {
                    if (lane == 0) atomicOr(gd.status, kStatusIterCap);
                    break;
                }
            }
        }
        const int32_t k = (int32_t)valid_res;
        if (lane < 4) o[4 * pk + j] = (p == psmall) ? k : n - k;
    }
    cursor_end(e, cur);
}

// ------------------------------------------------- gamma / Poisson VoP ---
// The reference has no GammaVGen/PoissonVGen. These follow its BinomialVGen
// design (Discrete.h:1124-1192): per-lane parameters in float, lockstep
// iterations, each lane keeps its first accepted candidate. Speculative wave
// evaluation as in op_binomial_vop.

// Gamma, Packet8f: three full draws per iteration (ux, vx, u3); lanes with
// alpha <= 1 run the GS step (NormalExp.h:536-558), others the tangent step
// (580-603) on the first two draws. Thread 8i+j evaluates iteration i of float
// lane j (u64 lane j/2, half j%2), eight iterations per wave step: one
// candidate per thread, and a lane's first accepted value reaches the thread
// that stores it through one bpermute. The per-lane constants are computed for
// eight packets at once (one packet-lane per thread) and fetched by bpermute.
__device__ void op_gamma_vop(Eng& e, uint64_t* carry, float* o, int64_t per, const float* alpha,
                             const float* beta, Guard gd, int lane) {
    constexpr uint64_t kLaneBits = 0x0101010101010101ull;  // threads of float lane 0
    const int64_t T = per / 8;
    if (T == 0) return;
    const int i = lane >> 3, j = lane & 7, l = j >> 1, h = j & 1;
    Cursor<4, 2> cur;
    cur.mt = e.mt;
    cur.carry = carry;
    cur.pos = e.idx;
    cur.nc = 0;
    cur.used = false;
    cur.eng = &e;
    float g_al = 0.f, g_be = 0.f, g_px = 0.f, g_inv = 0.f, g_sq = 0.f;
    for (int64_t pk = 0; pk < T; ++pk) {
        if ((pk & 7) == 0) {
            // constants of packets pk..pk+7: thread t holds packet pk + t/8, lane t%8
            const bool in = pk + i < T;
            g_al = in ? alpha[8 * pk + lane] : 1.0f;
            g_be = in ? beta[8 * pk + lane] : 1.0f;
            g_px = ERD_F(kE) / (g_al + ERD_F(kE));
            g_inv = 1.0f / g_al;
            g_sq = sqrtf(2.0f * g_al - 1.0f);
        }
        const int src = 8 * (int)(pk & 7) + j;
        const float al = __shfl(g_al, src), be = __shfl(g_be, src), px = __shfl(g_px, src),
                    inv = __shfl(g_inv, src), sq = __shfl(g_sq, src);
        const float am1 = al - 1.0f;
        const bool gs = al <= 1.0f;
        float res = 0.0f;
        uint32_t done = 0;  // bit q = float lane q has its value
        for (int it = 0;; it += 8) {
            cur.ensure(3, lane);
            const int ni_avail = cur.avail() / 3;
            const int ni = ni_avail < 8 ? ni_avail : 8;
            float xv = 0.f;
            bool acc = false;
            if (i < ni) {
                const uint64_t w0 = cur.word(3 * i, l), w1 = cur.word(3 * i + 1, l),
                               w2 = cur.word(3 * i + 2, l);
                const float u1 = ur_f32(h ? hi32(w0) : lo32(w0));
                const float u2 = ur_f32(h ? hi32(w1) : lo32(w1));
                if (gs) {
                    const float vx = u2 + 1.1920928955078125e-07f / 8;
                    const float lv = plog_pos(vx);
                    const float x1 = pexp(inv * lv);
                    const float q1 = pexp(-x1);
                    const float x2 = 1.0f - lv;
                    const float q2 = pexp(plog_pos(x2) * am1);
                    const bool c = u1 < px;
                    xv = c ? x1 : x2;
                    acc = ur_f32(h ? hi32(w2) : lo32(w2)) < (c ? q1 : q2);
                } else {
                    float ys, yc;
                    psincos_pos(ERD_F(kPi) * u1, ys, yc);
                    const float yx = ys / yc;
                    xv = sq * yx + am1;
                    const float ub = (yx * yx + 1.0f) * pexp(am1 * plog(xv / am1) - yx * sq);
                    acc = (0.0f < xv) && (u2 <= ub);
                }
            }
            const uint64_t A = __ballot(acc);
            int need = 0;
            bool all = true;
            uint32_t newly = 0;
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                if (done & (1u << q)) continue;
                const uint64_t m = A & (kLaneBits << q);
                if (m) {
                    const int first = (__ffsll((long long)m) - 1) >> 3;
                    need = need > first + 1 ? need : first + 1;
                    newly |= 1u << q;
                } else {
                    all = false;
                }
            }
            if (!all) need = ni;
            const uint64_t mj = A & (kLaneBits << j);
            const int src = mj ? (int)(__ffsll((long long)mj) - 1) : lane;
            const float got = __shfl(be * xv, src);
            if (newly & (1u << j)) res = got;
            done |= newly;
            cur.advance(3 * need);
            if (all) break;
            if (it > gd.lockstep_cap) {
                if (lane == 0) atomicOr(gd.status, kStatusIterCap);
                break;
            }
        }
        if (i == 0) o[8 * pk + j] = res;
    }
    __syncthreads();
    e.idx = cur.pos;
}

// Poisson, Packet4f via half draws: Knuth lockstep for lanes with mean < 12,
// then tangent lockstep for the others with float constants.
//
// The Knuth loop multiplies each small lane's running product by its uniform
// once per draw and stops at the first draw where no small lane's product
// still exceeds e^-mean; a lane's count is the number of draws before its own
// product fell to e^-mean. The products never increase (u < 1), so a lane's
// count is its own first stopping draw, and the packet consumes the largest
// of them plus one. Each lane's product chain therefore runs over a batch of
// up to 4 preloaded draws in its own thread, with one ballot per batch (most
// chains end within a batch: small means are mostly well below 1). The
// per-lane constants are computed for 16 packets at once (one packet-lane per
// thread) and fetched by bpermute; op_binomial_vop does the same.
__device__ void op_poisson_vop(Eng& e, uint64_t* carry, int32_t* o, int64_t per,
                               const float* mean, Guard gd, const float* lg, int lane) {
    const int64_t T = per / 4;
    if (T == 0) return;
    const int i = lane >> 2, j = lane & 3;
    HalfCursor cur = cursor_begin(e, carry, lane);
    const uint64_t jmask = 0x1111111111111111ull << j;
    float g_m = 0.f, g_nem = 0.f, g_psq = 0.f, g_plm = 0.f, g_pg1 = 0.f;
    for (int64_t pk = 0; pk < T; ++pk) {
        if ((pk & 15) == 0) {
            // constants of packets pk..pk+15: thread t holds packet pk + t/4, lane t%4
            g_m = pk + i < T ? mean[4 * pk + lane] : 1.0f;
            g_nem = pexp(-g_m);
            g_psq = sqrtf(2.0f * g_m);
            g_plm = plog(g_m);
            g_pg1 = g_m * g_plm - plgamma_approx(g_m + 1.0f);
        }
        const int src = 4 * (int)(pk & 15) + j;
        const float m = __shfl(g_m, src), nem = __shfl(g_nem, src), psq = __shfl(g_psq, src),
                    plm = __shfl(g_plm, src), pg1 = __shfl(g_pg1, src);
        const bool small = m < 12.0f;
        float res = 0.0f;
        const uint64_t V = __ballot(small) & 0xFull;
        if (V) {
            float val = 1.0f;
            int32_t first = small ? -1 : 0;  // draws before this lane stopped
            int32_t base = 0;                // draws consumed before the batch
            for (;;) {
                cur.ensure(1, lane);
                const int B = cur.avail() < 4 ? cur.avail() : 4;
                uint32_t wv[4];
#pragma unroll
                for (int t = 0; t < 4; ++t) wv[t] = t < B ? cur.word32(t, j) : 0u;
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    if (first < 0 && t < B) {
                        val = val * ur_f32(wv[t]);
                        if (!(nem < val)) first = base + t;
                    }
                }
                if (!(__ballot(first < 0) & 0xFull)) {
                    int32_t last = __builtin_amdgcn_readlane(first, 0);
#pragma unroll
                    for (int q = 1; q < 4; ++q) {
                        const int32_t f = __builtin_amdgcn_readlane(first, q);
                        last = last > f ? last : f;
                    }
                    cur.advance(last + 1 - base);
                    break;
                }
                cur.advance(B);
                base += B;
                if (base > gd.lockstep_cap) {
                    if (lane == 0) atomicOr(gd.status, kStatusIterCap);
                    first = first < 0 ? base : first;
                    break;
                }
            }
            res = (float)first;
        }
        if (V != 0xFull) {
            uint64_t done = V;
            for (int it = 0;; it += 16) {
                cur.ensure(2, lane);
                const int ni_avail = cur.avail() / 2;
                const int ni = ni_avail < 16 ? ni_avail : 16;
                float fres = 0.0f;
                bool fv = false;
                if (i < ni) {
                    float s, c;
                    psincos_pos(ERD_F(kPi) * ur_f32(cur.word32(2 * i, j)), s, c);
                    const float yx = s / c;
                    fres = truncf(psq * yx + m);
                    const float p1 = (yx * yx + 1.0f) * ERD_F(0.9);
                    const bool c1 = 0.0f <= fres;
                    const float lgf = c1 ? plgamma_int(lg, fres + 1.0f) : 0.0f;  // unused if !c1
                    const float p2 = pexp((fres * plm - lgf) - pg1);
                    const float u2 = ur_f32(cur.word32(2 * i + 1, j));
                    fv = c1 && (u2 <= p1 * p2);
                }
                const uint64_t A = __ballot(fv);
                int need = 0;
                bool all = true;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    if (done & (1ull << q)) continue;
                    const uint64_t mq = A & (0x1111111111111111ull << q);
                    if (mq) {
                        const int first = (__ffsll((long long)mq) - 1) >> 2;
                        need = need > first + 1 ? need : first + 1;
                    } else {
                        all = false;
                    }
                }
                if (!all) need = ni;
                // lane j's first accepted value, from the thread that drew it
                const uint64_t mine = A & jmask;
                const float got = __shfl(fres, mine ? (int)(__ffsll((long long)mine) - 1) : lane);
                if (!(done & (1ull << j)) && mine) res = got;
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    if (A & (0x1111111111111111ull << q)) done |= 1ull << q;
                cur.advance(2 * need);
                if (all) break;
                if (it > gd.lockstep_cap) {
                    if (lane == 0) atomicOr(gd.status, kStatusIterCap);
                    break;
                }
            }
        }
        if (lane < 4) o[4 * pk + j] = (int32_t)res;
    }
    cursor_end(e, cur);
}

// ------------------------------------------------------ discrete draws ---
// Lookup tables of DiscreteGen<int32_t, int32_t> (Discrete.h:352-485), built
// on the host: K < 16 uses the uint32 CDF (ncdf = K), otherwise the alias
// table (prob/alias of 2^bits entries, bitmask = 2^bits - 1).
struct DiscreteTable {
    const uint32_t* cdf;
    const int32_t* prob;
    const int32_t* alias;
    int32_t ncdf;      // K if the CDF path is used, else 0
    uint32_t bitmask;
};

// packetOp category of one 32-bit word (Discrete.h:458-475): CDF counts
// thresholds <= the 31-bit value; alias compares the logical word >> 1
// against prob (signed) and falls back to alias.
__device__ __forceinline__ int32_t discrete_packet(const DiscreteTable& t, uint32_t word) {
    if (t.ncdf) {
        const int32_t rx = (int32_t)(word & 0x7FFFFFFFu);
        int32_t ret = t.ncdf - 1;
        for (int i = 0; i + 1 < t.ncdf; ++i) ret -= (rx < (int32_t)t.cdf[i]) ? 1 : 0;
        return ret;
    }
    const int32_t albit = (int32_t)(word & t.bitmask);
    return ((int32_t)(word >> 1) < t.prob[albit]) ? albit : t.alias[albit];
}

// operator() category (Discrete.h:428-441): lower_bound over the CDF, or the
// alias test with an arithmetic shift and an unsigned compare.
__device__ __forceinline__ int32_t discrete_scalar(const DiscreteTable& t, uint32_t word) {
    const int32_t rx = (int32_t)word;
    if (t.ncdf) {
        const uint32_t r = (uint32_t)(rx & 0x7FFFFFFF);
        int32_t i = 0;
        while (i < t.ncdf - 1 && t.cdf[i] < r) ++i;
        return i;
    }
    const uint32_t albit = (uint32_t)rx & t.bitmask;
    const uint32_t alx = (uint32_t)(rx >> 1);
    return (alx < (uint32_t)t.prob[albit]) ? (int32_t)albit : t.alias[albit];
}

// DiscreteGen fill (packet path): one full draw gives 8 categories; the
// first Packet4i is returned, the second cached for the next packetOp
// (Discrete.h:443-483), so a unit = 1 draw = 8 outputs (a fill of 4 mod 8
// ends with an unused cached half).
struct DiscreteUnits {
    static constexpr int CALLS = 1;
    int32_t* out;
    DiscreteTable t;
    __device__ void rebase(int64_t off) { out += off; }
    __device__ static int64_t units(int64_t n) { return (n + 7) / 8; }
    __device__ void emit(int64_t per, int64_t u, int l, const uint64_t* w) const {
        const int32_t a = discrete_packet(t, lo32(w[0])), b = discrete_packet(t, hi32(w[0]));
        const int64_t i = 8 * u + 2 * l;  // words 2l, 2l+1 of the draw
        if (i < per) {
            out[i] = a;
            out[i + 1] = b;
        }
    }
};

// MultinomialGen::generate, trials < max(30(K-1), 100) (Multinomial.h:95-133):
// per sample, trials>>4 blocks of 16 draws (4 packetOps each), then the
// (trials&15)>>2 tail packetOps and trials&3 scalar draws (one full engine
// draw each, first word). The DiscreteGen cache (valid flag + cached upper
// half) carries from sample to sample. Counts go to out[s*K + category]
// (zeroed by the caller) with atomics.
__device__ void op_multinomial_small(Eng& e, int32_t* out, int64_t n, int32_t T, int32_t K,
                                     const DiscreteTable& tab, int lane) {
    int pos = e.idx;
    bool v = false;          // DiscreteGen cache valid
    uint64_t cw0 = 0, cw1 = 0;  // cached upper half (u64 lanes 2, 3), tempered
    const int32_t Q = 4 * (T >> 4) + ((T & 15) >> 2);
    const int32_t S = T & 3;
    for (int64_t s = 0; s < n; ++s) {
        int32_t* col = out + s * K;
        const int32_t nq = 4 * Q;  // packet draws of this sample
        int32_t first = 0;
        if (v && nq > 0) {
            if (lane < 4) {
                const uint64_t w = (lane >> 1) ? cw1 : cw0;
                atomicAdd(col + discrete_packet(tab, (lane & 1) ? hi32(w) : lo32(w)), 1);
            }
            first = 4;
        }
        // draws [first, nq) come from new full draws, 8 words each
        const int32_t nw = nq - first;
        const int32_t calls = (nw + 7) >> 3;
        int32_t done = 0;
        while (done < calls) {
            if (pos >= kN) {
                eng_refill(e, lane);
                pos = 0;
                e.twisted = true;
            }
            const int32_t take = (calls - done) < (kN - pos) ? (calls - done) : (kN - pos);
            for (int32_t q = lane; q < take * 8; q += 64) {
                const int32_t j = done * 8 + q;  // word index among the new draws
                if (j < nw) {
                    const uint64_t w = temper(e.mt[(pos + (q >> 3)) * kLanes + ((q & 7) >> 1)]);
                    atomicAdd(col + discrete_packet(tab, (q & 1) ? hi32(w) : lo32(w)), 1);
                }
            }
            pos += take;
            done += take;
        }
        if (nq > 0) {
            const bool v2 = v != ((Q & 1) != 0);
            if (v2) {
                cw0 = temper(e.mt[(pos - 1) * kLanes + 2]);
                cw1 = temper(e.mt[(pos - 1) * kLanes + 3]);
            }
            v = v2;
        }
        for (int32_t k = 0; k < S; ++k) {
            if (pos >= kN) {
                eng_refill(e, lane);
                pos = 0;
                e.twisted = true;
            }
            if (lane == 0) atomicAdd(col + discrete_scalar(tab, lo32(temper(e.mt[pos * kLanes]))), 1);
            pos += 1;
        }
    }
    __syncthreads();
    e.idx = pos;
}

// BinomialVGen<float>::packetOp at Packet8f (Discrete.h:1251-1330): full
// draws, lockstep Knuth for the lanes with mean < 1, then lockstep tangent
// rejection keeping each lane's first accepted candidate; thread 8i+j runs
// iteration i of lane j (8 iterations per step). T packets; the source S
// gives each packet's lane parameters and takes its results:
//   S::params(pk, lane, nt, p, psmall, plp, plq): lane j = lane & 7 of packet
//     pk (called by all 64 threads; nt = trials, psmall = pmin(p, 1 - p),
//     plp/plq = plog(psmall) / plog(1 - psmall));
//   S::result(pk, lane, nt, t): the lane's result t (called by all threads).
template <class S>
__device__ void op_binomial_v8(Eng& e, uint64_t* carry, int64_t T, S& src, Guard gd, const float* lg,
                               int lane) {
    if (T == 0) return;
    const int i = lane >> 3, j = lane & 7;
    Cursor<4, 1> cur;
    cur.mt = e.mt;
    cur.carry = carry;
    cur.pos = e.idx;
    cur.nc = 0;
    cur.used = false;
    cur.eng = &e;
    const uint64_t jmask = 0x0101010101010101ull << j;
    for (int64_t pk = 0; pk < T; ++pk) {
        int32_t nt;
        float p, psmall, plp, plq;
        src.params(pk, lane, nt, p, psmall, plp, plq);
        const float ptr = (float)nt;
        const float pmean = ptr * psmall;
        const float pg1 = plgamma_int(lg, ptr + 1.0f);  // trials + 1 is integer-valued
        const float psq = sqrtf((2.0f * pmean) * (1.0f - psmall));
        const bool valid = pmean < 1.0f;
        float valid_res = 0.0f;
        const uint64_t V = __ballot(valid) & 0xFFull;
        if (V) {
            const float nem = pexp(-pmean);
            float val = 1.0f, res = 0.0f;
            for (int it = 0;; ++it) {
                cur.ensure(1, lane);
                const float u = ur_f32(cur.word32(0, j));
                cur.advance(1);
                val = val * u;
                const bool c = (nem < val) && valid;
                if (!(__ballot(c) & 0xFFull)) break;
                res = res + (c ? 1.0f : 0.0f);
                if (it > gd.lockstep_cap) {
                    if (lane == 0) atomicOr(gd.status, kStatusIterCap);
                    break;
                }
            }
            valid_res = res;
        }
        if (V != 0xFFull) {
            uint64_t done = V;
            for (int it = 0;; it += 8) {
                cur.ensure(2, lane);
                const int ni_avail = cur.avail() / 2;
                const int ni = ni_avail < 8 ? ni_avail : 8;
                float fres = 0.0f;
                bool fv = false;
                if (i < ni) {
                    float s, c;
                    psincos_pos(ERD_F(kPi) * ur_f32(cur.word32(2 * i, j)), s, c);
                    const float ys = s / c;
                    fres = truncf(psq * ys + pmean);
                    const float pp1 = (ERD_F(1.2) * psq) * (1.0f + ys * ys);
                    const bool c1 = (0.0f <= fres) && (fres <= ptr);
                    float lg1 = 0.0f, lg2 = 0.0f;  // unused when the range test fails
                    if (c1) {
                        lg1 = plgamma_int(lg, fres + 1.0f);
                        lg2 = plgamma_int(lg, (ptr + 1.0f) - fres);
                    }
                    const float pp2 = pexp((((pg1 - lg1) - lg2) + fres * plp) + (ptr - fres) * plq);
                    const float u2 = ur_f32(cur.word32(2 * i + 1, j));
                    fv = c1 && (u2 <= pp1 * pp2);
                }
                const uint64_t A = __ballot(fv);
                int need = 0;
                bool all = true;
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    if (done & (1ull << q)) continue;
                    const uint64_t m = A & (0x0101010101010101ull << q);
                    if (m) {
                        const int fi = (__ffsll((long long)m) - 1) >> 3;
                        need = need > fi + 1 ? need : fi + 1;
                    } else {
                        all = false;
                    }
                }
                if (!all) need = ni;
                // lane j's first accepted value, from the thread that drew it
                const uint64_t mine = A & jmask;
                const float got = __shfl(fres, mine ? (int)(__ffsll((long long)mine) - 1) : lane);
                if (!(done & (1ull << j)) && mine) valid_res = got;
#pragma unroll
                for (int q = 0; q < 8; ++q)
                    if (A & (0x0101010101010101ull << q)) done |= 1ull << q;
                cur.advance(2 * need);
                if (all) break;
                if (it > gd.lockstep_cap) {
                    if (lane == 0) atomicOr(gd.status, kStatusIterCap);
                    break;
                }
            }
        }
        const int32_t k = (int32_t)valid_res;
        src.result(pk, lane, nt, (p == psmall) ? k : nt - k);
    }
    __syncthreads();
    e.idx = cur.pos;
}

// The per-row source of MultinomialGen's large-trials path
// (Multinomial.h:74-91): trials come from r[] (int32), p is one float for the
// row, and the row result t goes to out[i*stride] while r[i] -= t. REG
// (n <= 64): thread 8i+j keeps sample 8i+j's remaining trials in `rreg`
// instead of r[]. t0 >= 0: the first row, every sample has t0 trials. The row
// constants come from the caller, which computes them for many rows at once.
template <bool REG>
struct BinomialRowSrc {
    int32_t* out;
    int64_t stride;
    int32_t* r;
    int32_t& rreg;
    int32_t t0;
    float p, psmall, plp, plq;
    __device__ void params(int64_t pk, int lane, int32_t& nt, float& pp, float& ps, float& lp, float& lq) {
        const int j = lane & 7;
        nt = t0 >= 0 ? t0 : REG ? __shfl(rreg, 8 * (int)pk + j) : r[8 * pk + j];
        pp = p;
        ps = psmall;
        lp = plp;
        lq = plq;
    }
    __device__ void result(int64_t pk, int lane, int32_t nt, int32_t t) {
        const int i = lane >> 3, j = lane & 7;
        if (REG && i == pk) rreg = nt - t;
        if (lane < 8) {
            out[(8 * pk + j) * stride] = t;
            if (!REG) r[8 * pk + j] = nt - t;
        }
    }
};

template <bool REG>
__device__ void op_binomial_row8(Eng& e, uint64_t* carry, int32_t* out, int64_t stride, int32_t* r,
                                 int32_t& rreg, int32_t t0, int64_t n, float p, float psmall, float plp,
                                 float plq, Guard gd, const float* lg, int lane) {
    BinomialRowSrc<REG> src{out, stride, r, rreg, t0, p, psmall, plp, plq};
    op_binomial_v8(e, carry, n / 8, src, gd, lg, lane);
}

// binomial(urng, trials, p) with BinomialVGen<float> (Discrete.h:1196-1331,
// free functions 1824-1897): element k has trials[k] and p[k]; the
// reference carries the int32 trials and results as bit patterns in float
// lanes, here they are int32. The constants of 8 packets are computed at a
// time, thread t for element 8*pk + t, and read per packet with a shuffle.
struct BinomialVopFSrc {
    int32_t* out;
    Param<int32_t> trials;
    Param<float> prob;
    int64_t n;  // elements (8 per packet)
    int32_t g_n;
    float g_p, g_ps, g_lp, g_lq;
    __device__ void params(int64_t pk, int lane, int32_t& nt, float& pp, float& ps, float& lp, float& lq) {
        if ((pk & 7) == 0) {
            const bool in = 8 * pk + lane < n;
            g_n = in ? trials[8 * pk + lane] : 1;
            g_p = in ? prob[8 * pk + lane] : 0.5f;
            g_ps = (g_p < 1.0f - g_p) ? g_p : 1.0f - g_p;  // pmin(p, 1-p)
            g_lp = plog(g_ps);
            g_lq = plog(1.0f - g_ps);
        }
        const int src = 8 * (int)(pk & 7) + (lane & 7);
        nt = __shfl(g_n, src);
        pp = __shfl(g_p, src);
        ps = __shfl(g_ps, src);
        lp = __shfl(g_lp, src);
        lq = __shfl(g_lq, src);
    }
    __device__ void result(int64_t pk, int lane, int32_t, int32_t t) {
        if (lane < 8) out[8 * pk + lane] = t;
    }
};

}  // namespace erd
