// Batched multivariate kernels: FP64 Cholesky, split-BF16 MFMA transform
// (MvNormalGen's L*z, Wishart's chol*A), Wishart assembly and T*T^T, and the
// Dirichlet column normalisation.
//
// Instances of uneven dimension share one persistent launch: the host builds
// a tile list ordered by decreasing work and each workgroup pulls tiles from
// an atomic counter until the list is empty.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "erd_device.hpp"
#include "erd_mv.hpp"

namespace erd {

// ----------------------------------------------------------- Cholesky ---
// Right-looking Cholesky in FP64 of each instance's d x d matrix (column
// major, in place on a scratch copy; the strictly upper part is zeroed).
// info[i] = 0 on success, j+1 if the j-th pivot is not positive.
__global__ __launch_bounds__(256) void k_cholesky_f64(double* __restrict__ a,
                                                      const int64_t* __restrict__ off,
                                                      const int32_t* __restrict__ dims,
                                                      int32_t* __restrict__ info) {
    __shared__ double ljj_s;
    __shared__ int bad;
    const int inst = blockIdx.x;
    const int d = dims[inst];
    double* A = a + off[inst];
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    if (tid == 0) bad = 0;
    __syncthreads();
    for (int j = 0; j < d; ++j) {
        if (tid == 0) {
            const double ajj = A[(int64_t)j * d + j];
            if (!(ajj > 0.0)) {
                bad = j + 1;
                ljj_s = 1.0;
            } else {
                ljj_s = sqrt(ajj);
                A[(int64_t)j * d + j] = ljj_s;
            }
        }
        __syncthreads();
        if (bad) break;
        const double inv = 1.0 / ljj_s;
        for (int i = j + 1 + tid; i < d; i += 256) A[(int64_t)j * d + i] *= inv;
        __syncthreads();
        // A[i][k] -= L[i][j] * L[k][j] for j < k <= i: wave w takes columns
        // k = j+1+w, j+5+w, ...; lanes take consecutive rows (coalesced).
        for (int k = j + 1 + wave; k < d; k += 4) {
            const double lkj = A[(int64_t)j * d + k];
            for (int i = k + lane; i < d; i += 64)
                A[(int64_t)k * d + i] -= A[(int64_t)j * d + i] * lkj;
        }
        __syncthreads();
    }
    // zero the strictly upper triangle
    for (int64_t t = tid; t < (int64_t)d * d; t += 256) {
        const int col = (int)(t / d), row = (int)(t % d);
        if (row < col) A[t] = 0.0;
    }
    if (tid == 0) info[inst] = bad;
}

// Blocked left-looking Cholesky for d <= kCholBlkMaxDim: 32-column panels in
// LDS. For panel j0: load A[j0:d, j0:j0+32]; subtract L[j0:d, 0:j0] *
// L[j0:j0+32, 0:j0]^T (thread t takes panel row t with 32 FP64 accumulators;
// the 32 x 32 tiles of the other factor are staged in LDS); factor the
// tall-skinny panel in LDS (pivot, column scale, rank-1 updates of the panel's
// later columns); write it back with the strictly upper part zeroed. FP64
// throughout, the same factor as k_cholesky_f64 up to summation order.
// Dynamic LDS: (dmax * 32 + 32 * 32) doubles.
__global__ __launch_bounds__(256) void k_cholesky_f64_blk(double* __restrict__ a,
                                                          const int64_t* __restrict__ off,
                                                          const int32_t* __restrict__ dims,
                                                          int32_t* __restrict__ info, int dmax) {
    extern __shared__ double lds[];
    __shared__ double ljj_s;
    __shared__ int bad;
    constexpr int NB = 32;
    const int inst = blockIdx.x;
    const int d = dims[inst];
    double* A = a + off[inst];
    double* P = lds;                    // panel, column-major, leading dimension R
    double* B = lds + (int64_t)dmax * NB;  // [kk][c] tile of L[j0 + c][k0 + kk]
    const int tid = threadIdx.x;
    if (tid == 0) bad = 0;
    __syncthreads();
    for (int j0 = 0; j0 < d && !bad; j0 += NB) {
        const int w = min(NB, d - j0), R = d - j0;
        for (int t = tid; t < w * R; t += 256) {
            const int c = t / R, i = t - c * R;
            P[c * R + i] = A[(int64_t)(j0 + c) * d + j0 + i];
        }
        // update from the finished columns k < j0
        for (int i0 = 0; i0 < R; i0 += 256) {
            const int i = i0 + tid;
            const bool have = i < R;
            double acc[NB];
#pragma unroll
            for (int c = 0; c < NB; ++c) acc[c] = 0.0;
            for (int k0 = 0; k0 < j0; k0 += NB) {
                const int kw = min(NB, j0 - k0);
                __syncthreads();
                for (int t = tid; t < kw * NB; t += 256) {
                    const int kk = t / NB, c = t - kk * NB;
                    B[kk * NB + c] = c < w ? A[(int64_t)(k0 + kk) * d + j0 + c] : 0.0;
                }
                __syncthreads();
                for (int kk = 0; kk < kw; ++kk) {
                    const double av = have ? A[(int64_t)(k0 + kk) * d + j0 + i] : 0.0;
#pragma unroll
                    for (int c = 0; c < NB; ++c) acc[c] = fma(av, B[kk * NB + c], acc[c]);
                }
            }
            __syncthreads();
            if (have) {
#pragma unroll
                for (int c = 0; c < NB; ++c)
                    if (c < w) P[c * R + i] -= acc[c];
            }
        }
        __syncthreads();
        // factor the panel
        for (int c = 0; c < w; ++c) {
            if (tid == 0) {
                const double ajj = P[c * R + c];
                if (!(ajj > 0.0)) {
                    bad = j0 + c + 1;
                    ljj_s = 1.0;
                } else {
                    ljj_s = sqrt(ajj);
                    P[c * R + c] = ljj_s;
                }
            }
            __syncthreads();
            if (bad) break;
            const double inv = 1.0 / ljj_s;
            for (int i = c + 1 + tid; i < R; i += 256) P[c * R + i] *= inv;
            __syncthreads();
            const int rest = w - c - 1;
            for (int t = tid; t < rest * R; t += 256) {
                const int m = c + 1 + t / R, i = t % R;
                if (i >= m) P[m * R + i] -= P[c * R + i] * P[c * R + m];
            }
            __syncthreads();
        }
        if (bad) break;
        // write back: column j0 + c, rows j0 + c .. d-1 from the panel, rows above zeroed
        for (int t = tid; t < w * d; t += 256) {
            const int c = t / d, row = t - c * d;
            const int col = j0 + c;
            A[(int64_t)col * d + row] = row < col ? 0.0 : P[c * R + (row - j0)];
        }
        __syncthreads();
    }
    if (tid == 0) info[inst] = bad;
}

// -------------------------------------------------------- split-BF16 ---
__device__ __forceinline__ uint16_t bf16_rne(float x) {
    const uint32_t u = __float_as_uint(x);
    const uint32_t r = u + 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(r >> 16);
}
__device__ __forceinline__ float bf16_to_f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// Split of two floats into packed bf16 hi and lo words (element 0 in the low
// half): hi = bf16(x), lo = bf16(x - hi), both round-to-nearest-even by
// v_cvt_pk_bf16_f32, which equals bf16_rne bit for bit on finite inputs
// below the bf16 overflow threshold (all that L and Z hold).
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void split_bf16x2(float a, float b, uint32_t& hi, uint32_t& lo) {
    hi = __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2){a, b}, bf16x2));
    const float ra = a - __uint_as_float(hi << 16);
    const float rb = b - __uint_as_float(hi & 0xFFFF0000u);
    lo = __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2){ra, rb}, bf16x2));
}

// L (FP64, column-major d x d) -> row-major bf16 hi/lo planes of a padded
// dp x dp tile (dp = d rounded up to 16), plus the float L (column-major)
// that float consumers use. hi = bf16(l), lo = bf16(l - hi) with l = float(L).
__global__ void k_split_l(const double* __restrict__ l64, const int64_t* __restrict__ off,
                          const int32_t* __restrict__ dims, const int64_t* __restrict__ split_off,
                          uint16_t* __restrict__ hi, uint16_t* __restrict__ lo,
                          float* __restrict__ l32) {
    const int inst = blockIdx.y;
    const int d = dims[inst];
    const int dp = (d + 15) & ~15;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)dp * dp) return;
    const int r = (int)(t / dp), c = (int)(t % dp);  // row-major position
    float v = 0.0f;
    if (r < d && c < d) {
        v = (float)l64[off[inst] + (int64_t)c * d + r];
        l32[off[inst] + (int64_t)c * d + r] = v;
    }
    const uint16_t h = bf16_rne(v);
    hi[split_off[inst] + t] = h;
    lo[split_off[inst] + t] = bf16_rne(v - bf16_to_f(h));
}

// ------------------------------------------------------- MFMA L * Z ---
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
// Four floats at a 4-byte-aligned address (column starts are d floats
// apart, d arbitrary): one 16-byte global access.
typedef float f32x4_a4 __attribute__((ext_vector_type(4), aligned(4)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
// Global-memory views of the instance pointers: global (not flat)
// instructions, so outstanding loads do not hold up the LDS waits of the
// MFMA loop.
#define ERD_GLOBAL __attribute__((address_space(1)))
template <class T>
__device__ __forceinline__ ERD_GLOBAL T* gmem(T* p) {
    return (ERD_GLOBAL T*)p;
}

constexpr int KC = 32;       // k per chunk (one v_mfma_f32_16x16x32_bf16)
constexpr int KPAD = 40;     // LDS row pitch in bf16 (80 B: keeps 16 B alignment)

// Output tile = rows [row0, row0 + 16*MAXR) x NT = 64*NCB columns. Wave w
// owns column blocks w*NCB .. w*NCB+NCB-1 (16 columns each) and every 16-row
// block of the tile. Per 32-wide k chunk the workgroup
// stages L[:, k0:k0+32] (hi, lo) and Z[k0:k0+32, cols] (split on the fly)
// in LDS; each wave issues hi*hi + hi*lo + lo*hi MFMAs, skipping row blocks
// above the diagonal (L is lower triangular).
template <int NCB, int MAXR, bool BART>
__global__ __launch_bounds__(256, NCB == 4 || BART ? 2 : 3) void k_lz_gemm(const MvTile* __restrict__ tiles,
                                                 const int32_t* __restrict__ qbeg,
                                                 int32_t* __restrict__ counter,
                                                 const MvInstance* __restrict__ inst) {
    constexpr int NT = 64 * NCB;
    __shared__ __align__(16) uint16_t la[2][MAXR * 16 * KPAD];
    __shared__ __align__(16) uint16_t zb[2][NT * KPAD];
    __shared__ int tile_s;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int l16 = lane & 15, g = lane >> 4;
    // thread 0: the queue it takes tiles from, and how many it found empty
    int q = (int)(blockIdx.x % kLzQueues), emptied = 0;
    // thread 0: the next tile (-1: every queue empty), given the slot k
    // already taken from queue q
    auto next_tile = [&](int k) {
        while (emptied < kLzQueues) {
            if (k < qbeg[q + 1]) return k;
            q = q + 1 == kLzQueues ? 0 : q + 1;  // a queue found empty stays empty
            if (++emptied < kLzQueues) k = qbeg[q] + atomicAdd(counter + q, 1);
        }
        return -1;
    };
    if (tid == 0) tile_s = next_tile(qbeg[q] + atomicAdd(counter + q, 1));
    for (;;) {
        __syncthreads();
        const int t = tile_s;
        __syncthreads();
        if (t < 0) break;
        // the slot for the tile after this one is taken now and looked at
        // only when this tile is done, so the atomic's round trip overlaps
        // the tile's chunks
        int a_next = 0;
        if (tid == 0 && emptied < kLzQueues) a_next = atomicAdd(counter + q, 1);
        const MvTile tl = tiles[t];
        const MvInstance in = inst[tl.inst];
        const int d = in.d, dp = (d + 15) & ~15;
        const int row0 = tl.row0;
        const int R = (dp - row0) / 16 < MAXR ? (dp - row0) / 16 : MAXR;
        const int kmax = row0 + 16 * R;  // L[i][k] = 0 for k > i
        const int64_t n = in.n;
        const int64_t cend = tl.col0 + tl.ncols;  // the tile's columns are [c0, cend)
        const int64_t c0 = tl.col0;
        f32x4 acc[NCB][MAXR];
#pragma unroll
        for (int cb = 0; cb < NCB; ++cb)
#pragma unroll
            for (int rb = 0; rb < MAXR; ++rb) acc[cb][rb] = f32x4{0.f, 0.f, 0.f, 0.f};
        ERD_GLOBAL const uint16_t* Lh = gmem(in.lhi);
        ERD_GLOBAL const uint16_t* Ll = gmem(in.llo);
        ERD_GLOBAL const float* Z = gmem(in.z);
        ERD_GLOBAL const float* mean = in.mean ? gmem(in.mean) : nullptr;
        // The next chunk's L and Z are loaded into registers while the
        // current chunk's MFMAs run; the registers go to LDS (Z split into
        // hi/lo) between the two barriers of a step.
        constexpr int LQ = (MAXR * 16 * (KC / 8) + 255) / 256;  // L uint4 pairs per thread
        constexpr int ZQ = (NT * (KC / 8)) / 256;                // Z 8-float runs per thread
        static_assert((NT * (KC / 8)) % 256 == 0, "Z staging must cover the tile exactly");
        const int lq_n = R * 16 * (KC / 8);
        uint4 ph[LQ], pl[LQ];
        float pz[ZQ][8];
        // BART: the sample j and column i of each of this thread's Z columns,
        // and where the column's normals start in z (element k > i at z[zn + k])
        int zi[ZQ];
        int64_t zj[ZQ], zn[ZQ];
        if (BART) {
            const int64_t per_sample = (int64_t)d * (d - 1) / 2;
#pragma unroll
            for (int i = 0; i < ZQ; ++i) {
                const int c = (tid + 256 * i) / (KC / 8);
                const int64_t col = c0 + c < cend ? c0 + c : 0;
                zj[i] = col / d;
                zi[i] = (int)(col - zj[i] * d);
                zn[i] = zj[i] * per_sample + (int64_t)zi[i] * (d - 1) - (int64_t)zi[i] * (zi[i] - 1) / 2 - zi[i] - 1;
            }
        }
        auto fetch = [&](int k0) {
#pragma unroll
            for (int i = 0; i < LQ; ++i) {
                const int q = tid + 256 * i;
                ph[i] = uint4{0, 0, 0, 0};
                pl[i] = uint4{0, 0, 0, 0};
                if (q < lq_n) {
                    const int r = q / (KC / 8), kk = (q % (KC / 8)) * 8;
                    if (k0 + kk < dp) {
                        const u32x4 h = *reinterpret_cast<ERD_GLOBAL const u32x4*>(Lh + (int64_t)(row0 + r) * dp + k0 + kk);
                        const u32x4 l = *reinterpret_cast<ERD_GLOBAL const u32x4*>(Ll + (int64_t)(row0 + r) * dp + k0 + kk);
                        ph[i] = uint4{h[0], h[1], h[2], h[3]};
                        pl[i] = uint4{l[0], l[1], l[2], l[3]};
                    }
                }
            }
#pragma unroll
            for (int i = 0; i < ZQ; ++i) {
                const int q = tid + 256 * i;
                const int c = q / (KC / 8), kk = (q % (KC / 8)) * 8;
                const int64_t col = c0 + c;
                if (BART) {
                    const int kb = k0 + kk;
                    if (col < cend && kb > zi[i] && kb + 8 <= d) {
                        // all eight below the diagonal: consecutive normals
                        const f32x4_a4 a = *reinterpret_cast<ERD_GLOBAL const f32x4_a4*>(Z + zn[i] + kb);
                        const f32x4_a4 b = *reinterpret_cast<ERD_GLOBAL const f32x4_a4*>(Z + zn[i] + kb + 4);
#pragma unroll
                        for (int e = 0; e < 4; ++e) pz[i][e] = a[e], pz[i][4 + e] = b[e];
                    } else {
#pragma unroll
                        for (int e = 0; e < 8; ++e) {
                            const int k = kb + e;
                            float v = 0.0f;
                            if (col < cend && k < d) {
                                if (k == zi[i])
                                    v = sqrtf(in.chi[(int64_t)zi[i] * (n / d) + zj[i]]);
                                else if (k > zi[i])
                                    v = Z[zn[i] + k];
                            }
                            pz[i][e] = v;
                        }
                    }
                    continue;
                }
                if (col < cend && k0 + kk + 8 <= d) {
                    // 8 consecutive floats of column col: two 16-byte loads
                    const f32x4_a4 a = *reinterpret_cast<ERD_GLOBAL const f32x4_a4*>(Z + col * d + k0 + kk);
                    const f32x4_a4 b = *reinterpret_cast<ERD_GLOBAL const f32x4_a4*>(Z + col * d + k0 + kk + 4);
#pragma unroll
                    for (int e = 0; e < 4; ++e) pz[i][e] = a[e], pz[i][4 + e] = b[e];
                } else {
#pragma unroll
                    for (int e = 0; e < 8; ++e) {
                        const int k = k0 + kk + e;
                        pz[i][e] = (k < d && col < cend) ? Z[col * d + k] : 0.0f;
                    }
                }
            }
        };
        // k below the tile's first chunk contributes exact zeros (a Bartlett
        // column i is zero above row i)
        fetch(tl.kbeg);
        for (int k0 = tl.kbeg; k0 < kmax; k0 += KC) {
            // L[row0 + r, k0:k0+32] (row-major bf16) and Z[k0:k0+32, c0:c0+NT]
            // split into hi/lo, layout [col][k]
#pragma unroll
            for (int i = 0; i < LQ; ++i) {
                const int q = tid + 256 * i;
                if (q < lq_n) {
                    const int r = q / (KC / 8), kk = (q % (KC / 8)) * 8;
                    *reinterpret_cast<uint4*>(&la[0][r * KPAD + kk]) = ph[i];
                    *reinterpret_cast<uint4*>(&la[1][r * KPAD + kk]) = pl[i];
                }
            }
#pragma unroll
            for (int i = 0; i < ZQ; ++i) {
                const int q = tid + 256 * i;
                const int c = q / (KC / 8), kk = (q % (KC / 8)) * 8;
                uint4 vh, vl;
                split_bf16x2(pz[i][0], pz[i][1], vh.x, vl.x);
                split_bf16x2(pz[i][2], pz[i][3], vh.y, vl.y);
                split_bf16x2(pz[i][4], pz[i][5], vh.z, vl.z);
                split_bf16x2(pz[i][6], pz[i][7], vh.w, vl.w);
                *reinterpret_cast<uint4*>(&zb[0][c * KPAD + kk]) = vh;
                *reinterpret_cast<uint4*>(&zb[1][c * KPAD + kk]) = vl;
            }
            __syncthreads();
            if (k0 + KC < kmax) fetch(k0 + KC);
#pragma unroll
            for (int cb = 0; cb < NCB; ++cb) {
                const int col = (wave * NCB + cb) * 16 + l16;
                const bf16x8 bh = *reinterpret_cast<const bf16x8*>(&zb[0][col * KPAD + 8 * g]);
                const bf16x8 bl = *reinterpret_cast<const bf16x8*>(&zb[1][col * KPAD + 8 * g]);
#pragma unroll
                for (int rb = 0; rb < MAXR; ++rb) {
                    if (rb < R && row0 + rb * 16 + 15 >= k0) {
                        const int row = rb * 16 + l16;
                        const bf16x8 ah = *reinterpret_cast<const bf16x8*>(&la[0][row * KPAD + 8 * g]);
                        const bf16x8 al = *reinterpret_cast<const bf16x8*>(&la[1][row * KPAD + 8 * g]);
                        acc[cb][rb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, acc[cb][rb], 0, 0, 0);
                        acc[cb][rb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl, acc[cb][rb], 0, 0, 0);
                        acc[cb][rb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, acc[cb][rb], 0, 0, 0);
                    }
                }
            }
            __syncthreads();
        }
        // epilogue: X[i][j] = acc + mean[i]; lane holds rows 4g..4g+3 of column l16
        ERD_GLOBAL float* X = gmem(in.x);
#pragma unroll
        for (int cb = 0; cb < NCB; ++cb) {
            const int64_t col = c0 + (wave * NCB + cb) * 16 + l16;
            if (col >= cend) continue;
#pragma unroll
            for (int rb = 0; rb < MAXR; ++rb) {
                if (rb >= R) continue;
                const int rowb = row0 + rb * 16 + 4 * g;
                if (rowb + 4 <= d) {
                    // rows rowb..rowb+3 of column col: one 16-byte store
                    f32x4_a4 v;
#pragma unroll
                    for (int r = 0; r < 4; ++r) v[r] = acc[cb][rb][r] + (mean ? mean[rowb + r] : 0.0f);
                    *reinterpret_cast<ERD_GLOBAL f32x4_a4*>(X + col * d + rowb) = v;
                } else {
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int row = rowb + r;
                        if (row < d) {
                            const float mu = mean ? mean[row] : 0.0f;
                            X[col * d + row] = acc[cb][rb][r] + mu;
                        }
                    }
                }
            }
        }
        if (tid == 0) tile_s = next_tile(emptied < kLzQueues ? qbeg[q] + a_next : 0);
    }
}

int lz_blocks_per_cu(int variant) {
    static int cache[kLzVariants] = {};
    if (variant < 0 || variant >= kLzVariants) return 1;
    if (cache[variant] == 0) {
        const void* f = variant == 0   ? (const void*)k_lz_gemm<1, 16, false>
                        : variant == 1 ? (const void*)k_lz_gemm<2, 8, false>
                        : variant == 2 ? (const void*)k_lz_gemm<4, 4, false>
                                       : (const void*)k_lz_gemm<1, 4, false>;
        int b = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, f, 256, 0) != hipSuccess || b < 1) b = 2;
        cache[variant] = b;
    }
    return cache[variant];
}

template <bool BART>
static void launch_lz_variant(int variant, int grid, const MvTile* tiles, const int32_t* qbeg, int32_t* counter,
                              const MvInstance* inst, hipStream_t s) {
    switch (variant) {
        case 0: hipLaunchKernelGGL((k_lz_gemm<1, 16, BART>), dim3(grid), dim3(256), 0, s, tiles, qbeg, counter, inst); break;
        case 1: hipLaunchKernelGGL((k_lz_gemm<2, 8, BART>), dim3(grid), dim3(256), 0, s, tiles, qbeg, counter, inst); break;
        case 2: hipLaunchKernelGGL((k_lz_gemm<4, 4, BART>), dim3(grid), dim3(256), 0, s, tiles, qbeg, counter, inst); break;
        default: hipLaunchKernelGGL((k_lz_gemm<1, 4, BART>), dim3(grid), dim3(256), 0, s, tiles, qbeg, counter, inst); break;
    }
}

hipError_t launch_lz_gemm(int variant, int grid, const MvTile* tiles, const int32_t* qbeg, int32_t* counter,
                          const MvInstance* inst, hipStream_t s, bool bartlett) {
    if (grid <= 0) return hipSuccess;
    if (bartlett)
        launch_lz_variant<true>(variant, grid, tiles, qbeg, counter, inst, s);
    else
        launch_lz_variant<false>(variant, grid, tiles, qbeg, counter, inst, s);
    return hipGetLastError();
}

hipError_t launch_cholesky_f64(double* a, const int64_t* off, const int32_t* dims, int32_t* info,
                               int n_inst, int dmax, hipStream_t s) {
    if (n_inst <= 0) return hipSuccess;
    if (dmax <= kCholBlkMaxDim) {
        const size_t lds = sizeof(double) * ((size_t)dmax * 32 + 32 * 32);
        hipError_t err = hipFuncSetAttribute((const void*)k_cholesky_f64_blk,
                                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (err != hipSuccess) return err;
        hipLaunchKernelGGL(k_cholesky_f64_blk, dim3(n_inst), dim3(256), lds, s, a, off, dims, info, dmax);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(k_cholesky_f64, dim3(n_inst), dim3(256), 0, s, a, off, dims, info);
    return hipGetLastError();
}

hipError_t launch_split_l(const double* l64, const int64_t* off, const int32_t* dims,
                          const int64_t* split_off, uint16_t* hi, uint16_t* lo, float* l32, int n_inst,
                          int max_dp, hipStream_t s) {
    if (n_inst <= 0) return hipSuccess;
    const int64_t elems = (int64_t)max_dp * max_dp;
    dim3 grid((unsigned)((elems + 255) / 256), (unsigned)n_inst);
    hipLaunchKernelGGL(k_split_l, grid, dim3(256), 0, s, l64, off, dims, split_off, hi, lo, l32);
    return hipGetLastError();
}

}  // namespace erd

namespace erd {

// --------------------------------------------------------------- Wishart ---
// W_j = T_j T_j^T (WishartGen, MvNormal.h:271-275) or U_j^T U_j (U = T^-1,
// InvWishartGen, MvNormal.h:433-439) in FP32, per sample, over 64 x 64 output tiles with
// ta >= tb (tile units of 64). Per 32-wide k chunk the workgroup stages the
// tile's 64 rows of both operands in LDS ([k][row]); each wave accumulates a
// 32 x 32 quadrant with FP32 MFMA (v_mfma_f32_32x32x2_f32, FP32 products and
// sums), over k from 0 (Wishart: T[b][k] = 0 beyond k = b) or from the tile's
// first row (inverse: U[k][a] = 0 for k < a); the extra terms the tile
// bounds add are exact zeros. Both halves of W are written from the same
// accumulator, so W is exactly symmetric; diagonal tiles write only their
// lower half.
constexpr int kSyT = 64, kSyK = 32, kSyPitch = kSyT + 4;

template <bool UTU>
__global__ __launch_bounds__(256) void k_syrk_tt(const SyrkTile* __restrict__ tiles,
                                                 const MvInstance* __restrict__ inst,
                                                 float* __restrict__ w, const int64_t* __restrict__ w_off) {
    __shared__ __align__(16) float sm[2][kSyK][kSyPitch];
    float(&ta_s)[kSyK][kSyPitch] = sm[0];
    float(&tb_s)[kSyK][kSyPitch] = sm[1];
    const SyrkTile tl = tiles[blockIdx.x];
    if (tl.inst < 0) return;  // padding of the per-XCD tile order
    const MvInstance in = inst[tl.inst];
    const int d = in.d;
    const float* T = in.x + (int64_t)tl.sample * d * d;  // d x d column-major
    float* W = w + w_off[tl.inst] + (int64_t)tl.sample * d * d;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // wave w computes the 32 x 32 quadrant (rows qa.., columns qb..) of the tile
    const int qa = (wave >> 1) * 32, qb = (wave & 1) * 32;
    const int a0 = tl.ta * kSyT, b0 = tl.tb * kSyT;
    f32x16 acc;
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[r] = 0.0f;
    const int kbeg = UTU ? a0 : 0;
    const int kend = UTU ? d : (b0 + kSyT < d ? b0 + kSyT : d);
    // element q = tid + 256 i of a chunk's staging: row r, k offset kk
    constexpr int kSyQ = kSyT * kSyK / 256;
    auto rk = [&](int i, int& r, int& kk) {
        const int q = tid + 256 * i;
        if (UTU) {  // consecutive threads walk down a column of U
            kk = q & (kSyK - 1);
            r = q / kSyK;
        } else {    // consecutive threads walk down a column of T
            r = q & (kSyT - 1);
            kk = q / kSyT;
        }
    };
    // the next chunk's operands are loaded into registers while the current
    // chunk's products run
    float pa[kSyQ], pb[kSyQ];
    auto fetch = [&](int k0) {
#pragma unroll
        for (int i = 0; i < kSyQ; ++i) {
            int r, kk;
            rk(i, r, kk);
            const int k = k0 + kk;
            pa[i] = 0.0f;
            pb[i] = 0.0f;
            if (k < d) {
                if (a0 + r < d) pa[i] = UTU ? T[(int64_t)(a0 + r) * d + k] : T[(int64_t)k * d + a0 + r];
                if (b0 + r < d) pb[i] = UTU ? T[(int64_t)(b0 + r) * d + k] : T[(int64_t)k * d + b0 + r];
            }
        }
    };
    if (kbeg < kend) fetch(kbeg);
    for (int k0 = kbeg; k0 < kend; k0 += kSyK) {
#pragma unroll
        for (int i = 0; i < kSyQ; ++i) {
            int r, kk;
            rk(i, r, kk);
            ta_s[kk][r] = pa[i];
            tb_s[kk][r] = pb[i];
        }
        __syncthreads();
        if (k0 + kSyK < kend) fetch(k0 + kSyK);
        // D[i][j] += sum over k of A[i][k] B[k][j] on the matrix cores, two k per
        // instruction: lane l holds A[l % 32][k + l / 32] and B[k + l / 32][l % 32]
#pragma unroll
        for (int kk = 0; kk < kSyK; kk += 2) {
            const float av = ta_s[kk + (lane >> 5)][qa + (lane & 31)];
            const float bv = tb_s[kk + (lane >> 5)][qb + (lane & 31)];
            acc = __builtin_amdgcn_mfma_f32_32x32x2f32(av, bv, acc, 0, 0, 0);
        }
        __syncthreads();
    }
    // the tile through LDS (the operand buffers are free after the last
    // barrier), so that both halves of W are written along columns
    static_assert(sizeof(sm) >= sizeof(float) * kSyT * (kSyT + 1), "the tile fits the operand LDS");
    float(*ot)[kSyT + 1] = reinterpret_cast<float(*)[kSyT + 1]>(&sm[0][0][0]);
    // D[i][j] of the quadrant is register r of lane l with j = l % 32,
    // i = 8 (r / 4) + 4 (l / 32) + r % 4
#pragma unroll
    for (int r = 0; r < 16; ++r) ot[qa + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3)][qb + (lane & 31)] = acc[r];
    __syncthreads();
    const bool diag = tl.ta == tl.tb;
    const int c = tid & (kSyT - 1), r0 = tid >> 6;  // 64 consecutive threads along a column
    for (int r = r0; r < kSyT; r += 4) {
        // column b0 + r of W, rows a0 + c: W(a, b) = ot[a - a0][b - b0]
        const int ra = a0 + c, rb = b0 + r;
        if (ra < d && rb < d && (!diag || ra >= rb)) W[(int64_t)rb * d + ra] = ot[c][r];
        // column a0 + r of W, rows b0 + c: W(b, a) = ot[a - a0][b - b0]
        const int sa = a0 + r, sb = b0 + c;
        if (sa < d && sb < d && (!diag || sa >= sb)) W[(int64_t)sa * d + sb] = ot[r][c];
    }
}

// ------------------------------------------------------------- Dirichlet ---
// ret.array().rowwise() /= ret.array().colwise().sum() (Multinomial.h:200-201)
// on the K x n result; g holds row k of gamma draws contiguously (g[k*n+s]),
// out is K x n column-major (out[s*K+k]). One wave per 64 samples: each lane
// sums its sample's column in category order (coalesced over samples), then
// 64x64 tiles are divided and transposed through LDS so that the stores run
// along categories.
constexpr int kDirTile = 64;

__global__ __launch_bounds__(64) void k_dirichlet_norm(const int32_t* __restrict__ dims,
                                                       const int64_t* __restrict__ samples,
                                                       const int64_t* __restrict__ g_off,
                                                       const float* __restrict__ g, float* __restrict__ out) {
    __shared__ float tile[kDirTile][kDirTile + 1];
    const int i_inst = blockIdx.y;
    const int K = dims[i_inst];
    const int64_t n = samples[i_inst];
    const int64_t s0 = (int64_t)blockIdx.x * kDirTile;
    if (s0 >= n) return;
    const int lane = threadIdx.x;
    const float* G = g + g_off[i_inst];
    float* O = out + g_off[i_inst];
    const int64_t s = s0 + lane;
    const int ns = (int)min<int64_t>(kDirTile, n - s0);
    float sum = 0.0f;
    // the loads are unrolled ahead of the adds, which stay in category order
    if (lane < ns) {
#pragma unroll 8
        for (int k = 0; k < K; ++k) sum += G[(int64_t)k * n + s];
    }
    for (int k0 = 0; k0 < K; k0 += kDirTile) {
        const int nk = min(kDirTile, K - k0);
        if (lane < ns) {
#pragma unroll 8
            for (int i = 0; i < nk; ++i) tile[i][lane] = G[(int64_t)(k0 + i) * n + s] / sum;
        }
        __syncthreads();
        if (lane < nk)
            for (int j = 0; j < ns; ++j) O[(s0 + j) * K + k0 + lane] = tile[lane][j];
        __syncthreads();
    }
}

hipError_t launch_syrk_tt(const SyrkTile* tiles, int n_tiles, const MvInstance* inst, float* w,
                          const int64_t* w_off, hipStream_t s, bool utu) {
    if (n_tiles <= 0) return hipSuccess;
    if (utu)
        hipLaunchKernelGGL(k_syrk_tt<true>, dim3((unsigned)n_tiles), dim3(256), 0, s, tiles, inst, w, w_off);
    else
        hipLaunchKernelGGL(k_syrk_tt<false>, dim3((unsigned)n_tiles), dim3(256), 0, s, tiles, inst, w, w_off);
    return hipGetLastError();
}

// -------------------------------------------------------- inverse Wishart ---
// Inverse of an SPD matrix from its FP64 Cholesky factor L (column-major,
// upper part zero): S^-1 = L^-T L^-1, for InvWishartGen's get_lt(scale.
// inverse()) (MvNormal.h:375-380). One workgroup per instance. Thread c owns
// column c of L^-1 (forward substitution, rows in order; every thread reads
// the same L[i][k] at a step), stored row-major in `w` so that a step's
// accesses are contiguous; then S^-1[i][j] = sum_{k >= max(i,j)} L^-1[k][i]
// L^-1[k][j], written column-major into `out`.
__global__ __launch_bounds__(256) void k_spd_inverse_f64(const double* __restrict__ l, double* __restrict__ w,
                                                        double* __restrict__ out, const int64_t* __restrict__ off,
                                                        const int32_t* __restrict__ dims) {
    const int inst = blockIdx.x;
    const int d = dims[inst];
    const double* L = l + off[inst];
    double* V = w + off[inst];  // V[k*d + c] = L^-1[k][c]
    double* S = out + off[inst];
    for (int c = threadIdx.x; c < d; c += blockDim.x) {
        for (int i = 0; i < d; ++i) {
            double v = 0.0;
            if (i == c) {
                v = 1.0 / L[(int64_t)i * d + i];
            } else if (i > c) {
                double acc = 0.0;
                for (int k = c; k < i; ++k) acc += L[(int64_t)k * d + i] * V[(int64_t)k * d + c];
                v = -acc / L[(int64_t)i * d + i];
            }
            V[(int64_t)i * d + c] = v;
        }
    }
    __syncthreads();
    for (int j = threadIdx.x; j < d; j += blockDim.x) {
        for (int i = 0; i < d; ++i) {
            double acc = 0.0;
            for (int k = i > j ? i : j; k < d; ++k) acc += V[(int64_t)k * d + i] * V[(int64_t)k * d + j];
            S[(int64_t)j * d + i] = acc;
        }
    }
}

// U = T^-1 for the lower-triangular T of each InvWishart sample
// (MvNormal.h:433-438: t.triangularView<Lower>().solve(identity)), in FP32,
// in two launches over the (instance, sample, 32-column panel J) list:
// k_trtri_diag inverts every diagonal block, U_JJ = T_JJ^-1 (forward
// substitution, one column per thread, independent across blocks), straight
// into U; k_trtri_panel then forms each panel by 32-row blocks I > J,
// X_I = U_II (E_IJ - sum_{K=J}^{I-1} T_IK X_K), with the panel X (d x 32) in
// LDS: block products only, no serial solve inside a panel. Rows above the
// panel are zero.

// U_JJ = T_JJ^-1 of panel tile t, eight tiles per workgroup (32 threads each,
// thread c: column c). Diagonal blocks past d are identity-padded.
__global__ __launch_bounds__(256) void k_trtri_diag(const SyrkTile* __restrict__ tiles, int n_tiles,
                                                    const MvInstance* __restrict__ inst, float* __restrict__ u_out,
                                                    const int64_t* __restrict__ u_off) {
    __shared__ float Tb[8][32][33];
    const int grp = threadIdx.x >> 5, c = threadIdx.x & 31;
    const int t = blockIdx.x * 8 + grp;
    SyrkTile tl{-1, 0, 0, 0};
    if (t < n_tiles) tl = tiles[t];
    const bool live = tl.inst >= 0;  // padding of the per-XCD panel order
    int d = 0, J = 0;
    const float* T = nullptr;
    float* U = nullptr;
    if (live) {
        const MvInstance in = inst[tl.inst];
        d = in.d;
        J = tl.ta;
        T = in.x + (int64_t)tl.sample * d * d;  // d x d column-major
        U = u_out + u_off[tl.inst] + (int64_t)tl.sample * d * d;
        // column by column, thread c reading row c: 128 contiguous bytes a column
        const int row = J * 32 + c;
        for (int k = 0; k < 32; ++k) {
            const int col = J * 32 + k;
            Tb[grp][c][k] = (row < d && col < d) ? T[(int64_t)col * d + row] : (c == k ? 1.0f : 0.0f);
        }
    }
    __syncthreads();
    float v[32];
#pragma unroll
    for (int r = 0; r < 32; ++r) v[r] = r == c ? 1.0f : 0.0f;
    if (live) {
#pragma unroll
        for (int r = 0; r < 32; ++r) {
            float x = v[r];
#pragma unroll
            for (int k = 0; k < r; ++k) x -= Tb[grp][r][k] * v[k];
            v[r] = x / Tb[grp][r][r];
        }
    }
    // back through LDS, so that the stores also run down the columns
    __syncthreads();
    if (live) {
#pragma unroll
        for (int r = 0; r < 32; ++r) Tb[grp][r][c] = v[r];
    }
    __syncthreads();
    if (!live) return;
    const int row = J * 32 + c;
    if (row < d) {
        for (int k = 0; k < 32; ++k) {
            const int col = J * 32 + k;
            if (col < d) U[(int64_t)col * d + row] = Tb[grp][c][k];
        }
    }
}

__global__ __launch_bounds__(256) void k_trtri_panel(const SyrkTile* __restrict__ tiles,
                                                     const MvInstance* __restrict__ inst, float* __restrict__ u_out,
                                                     const int64_t* __restrict__ u_off) {
    // X: the panel's rows J*32 .. nb*32-1 (row-relative), 33 floats a row;
    // dynamic LDS sized for the launch's largest nb - J
    extern __shared__ float X[];
    __shared__ float Tb[2][32][33];
    const SyrkTile tl = tiles[blockIdx.x];  // ta = panel J
    if (tl.inst < 0) return;               // padding of the per-XCD panel order
    const MvInstance in = inst[tl.inst];
    const int d = in.d;
    const float* T = in.x + (int64_t)tl.sample * d * d;  // d x d column-major
    float* U = u_out + u_off[tl.inst] + (int64_t)tl.sample * d * d;
    const int J = tl.ta, nb = (d + 31) / 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // column tx, rows ty*4..ty*4+3
    // Tiles are used in the order (I, K) = (J, J), (J+1, J), (J+1, J+1), ...:
    // for each block row I the updates T_IK, K = J..I-1, then U_II (the
    // diagonal block of U, from k_trtri_diag; identity-padded past d). The
    // next tile is fetched into registers while the current one is in use;
    // Tb is double-buffered, so one barrier per tile keeps writers and
    // readers apart.
    float pre[4];
    auto fetch = [&](int I, int K) {
        const float* src = I == K ? U : T;
#pragma unroll
        for (int q4 = 0; q4 < 4; ++q4) {
            const int q = threadIdx.x + 256 * q4;
            const int r = q & 31, kk = q >> 5;
            const int row = I * 32 + r, col = K * 32 + kk;
            pre[q4] = (row < d && col < d) ? src[(int64_t)col * d + row] : ((I == K && r == kk) ? 1.0f : 0.0f);
        }
    };
    fetch(J, J);
    int buf = 0;
    for (int I = J; I < nb; ++I) {
        float acc[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[q] = (I == J && ty * 4 + q == tx) ? 1.0f : 0.0f;
        for (int K = J; K <= I; ++K) {
            float(*tb)[33] = Tb[buf];
#pragma unroll
            for (int q4 = 0; q4 < 4; ++q4) {
                const int q = threadIdx.x + 256 * q4;
                tb[q & 31][q >> 5] = pre[q4];
            }
            if (K == I) {
#pragma unroll
                for (int q = 0; q < 4; ++q) X[((I - J) * 32 + ty * 4 + q) * 33 + tx] = acc[q];
            }
            __syncthreads();
            const int nI = K < I ? I : I + 1, nK = K < I ? K + 1 : J;
            if (nI < nb) fetch(nI, nK);
            if (K < I) {
#pragma unroll 8
                for (int kk = 0; kk < 32; ++kk) {
                    const float xv = X[((K - J) * 32 + kk) * 33 + tx];
#pragma unroll
                    for (int q = 0; q < 4; ++q) acc[q] = fmaf(-tb[ty * 4 + q][kk], xv, acc[q]);
                }
            } else {
                // X_I = U_II R_I (U_II lower triangular; R_I = E_JJ for I = J,
                // where X_J = U_JJ itself)
                float o[4];
                if (I == J) {
#pragma unroll
                    for (int q = 0; q < 4; ++q) o[q] = tb[ty * 4 + q][tx];
                } else {
#pragma unroll
                    for (int q = 0; q < 4; ++q) o[q] = 0.0f;
#pragma unroll 8
                    for (int kk = 0; kk < 32; ++kk) {
                        const float xv = X[((I - J) * 32 + kk) * 33 + tx];
#pragma unroll
                        for (int q = 0; q < 4; ++q) o[q] = fmaf(tb[ty * 4 + q][kk], xv, o[q]);
                    }
                }
                __syncthreads();
#pragma unroll
                for (int q = 0; q < 4; ++q) X[((I - J) * 32 + ty * 4 + q) * 33 + tx] = o[q];
                __syncthreads();
            }
            buf ^= 1;
        }
    }
    // store the panel's columns (column-major U; zeros above the panel's rows)
    // (the diagonal block, written by k_trtri_diag, is left as it is: other
    // panels read it)
    for (int q = threadIdx.x; q < 32 * d; q += 256) {
        const int row = q % d, c = q / d;
        const int col = J * 32 + c;
        if (col < d && (row < J * 32 || row >= J * 32 + 32))
            U[(int64_t)col * d + row] = row < J * 32 ? 0.0f : X[(row - J * 32) * 33 + c];
    }
}

hipError_t launch_spd_inverse_f64(const double* l, double* w, double* out, const int64_t* off, const int32_t* dims,
                                  int n_inst, hipStream_t s) {
    if (n_inst <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_spd_inverse_f64, dim3((unsigned)n_inst), dim3(256), 0, s, l, w, out, off, dims);
    return hipGetLastError();
}

hipError_t launch_trtri_panel(const SyrkTile* tiles, int n_tiles, int max_rows, const MvInstance* inst, float* u,
                              const int64_t* u_off, hipStream_t s) {
    if (n_tiles <= 0) return hipSuccess;
    if (max_rows <= 0 || max_rows > kTrtriMaxDim || max_rows % 32) return hipErrorInvalidValue;
    const size_t lds = sizeof(float) * (size_t)max_rows * 33;
    hipError_t err = hipFuncSetAttribute((const void*)k_trtri_panel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                         (int)lds);
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL(k_trtri_diag, dim3((unsigned)((n_tiles + 7) / 8)), dim3(256), 0, s, tiles, n_tiles, inst, u,
                       u_off);
    hipLaunchKernelGGL(k_trtri_panel, dim3((unsigned)n_tiles), dim3(256), lds, s, tiles, inst, u, u_off);
    return hipGetLastError();
}

hipError_t launch_dirichlet_norm(const int32_t* dims, const int64_t* samples, const int64_t* g_off,
                                 const float* g, float* out, int n_inst, int64_t max_n, hipStream_t s) {
    if (n_inst <= 0) return hipSuccess;
    dim3 grid((unsigned)((max_n + kDirTile - 1) / kDirTile), (unsigned)n_inst);
    hipLaunchKernelGGL(k_dirichlet_norm, grid, dim3(kDirTile), 0, s, dims, samples, g_off, g, out);
    return hipGetLastError();
}

}  // namespace erd
