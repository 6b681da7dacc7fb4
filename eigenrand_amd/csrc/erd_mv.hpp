// Shared host/device descriptors of the batched multivariate kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace erd {

// One instance of a batched L*Z (+ mean) transform.
struct MvInstance {
    int32_t d;             // dimension
    int32_t pad;
    int64_t n;             // columns of Z / X
    const uint16_t* lhi;   // padded dp x dp row-major bf16 planes of L (hi, lo)
    const uint16_t* llo;
    const float* z;        // d x n, column-major
    float* x;              // d x n, column-major
    const float* mean;     // d floats, or null
    // Wishart's Bartlett matrices built in the product's operand fetch (no A
    // in memory): with chi set, column i of sample j of Z (n = d * samples
    // columns) is 0 above row i, sqrt(chi[i * samples + j]) at row i, and
    // below it the next normals of z (sample j's d(d-1)/2 normals, column by
    // column), as WishartGen's rand_mat (MvNormal.h:240-266)
    const float* chi;
};

// A work item of the persistent transform: rows [row0, row0 + 16*MAXR) x
// columns [col0, col0 + NT) of one instance.
struct MvTile {
    int32_t inst;
    int32_t row0;
    int64_t col0;
    int32_t kbeg;   // first k chunk (a multiple of 32)
    int32_t ncols;  // columns of the tile: [col0, col0 + ncols)
};

// FP64 Cholesky of n instances; dmax = the largest d (host value): up to
// kCholBlkMaxDim the blocked LDS-panel kernel, above it the unblocked one.
constexpr int kCholBlkMaxDim = 512;
hipError_t launch_cholesky_f64(double* a, const int64_t* off, const int32_t* dims, int32_t* info,
                               int n_inst, int dmax, hipStream_t s);
hipError_t launch_split_l(const double* l64, const int64_t* off, const int32_t* dims,
                          const int64_t* split_off, uint16_t* hi, uint16_t* lo, float* l32, int n_inst,
                          int max_dp, hipStream_t s);
// variant: 0 = (1 col block per wave, 256-row tiles), 1 = (2, 128-row), 2 = (4, 64-row),
// 3 = (1, 64-row). Any dimension works with any variant (rows are tiled).
// The tiles come in kLzQueues queues, tiles[qbeg[q], qbeg[q+1]); workgroup b
// serves queue b % kLzQueues first (the XCD it is dispatched to: workgroups
// go to the 8 XCDs round robin), then takes what is left in the others.
// counter: kLzQueues zeroed ints.
constexpr int kLzQueues = 8;
// Workgroups of L*Z variant v resident per CU (from its register and LDS use).
int lz_blocks_per_cu(int variant);
hipError_t launch_lz_gemm(int variant, int grid, const MvTile* tiles, const int32_t* qbeg, int32_t* counter,
                          const MvInstance* inst, hipStream_t s, bool bartlett = false);

struct SyrkTile {
    int32_t inst;
    int32_t sample;
    int32_t ta, tb;
};
// W = T T^T per sample (Wishart), or W = U^T U with U lower triangular (utu).
constexpr int kXcds = 8;  // XCDs of the device; workgroups are dispatched to them round-robin
constexpr int kSyrkTile = 64;  // output tile edge of launch_syrk_tt (SyrkTile ta, tb in these units)
hipError_t launch_syrk_tt(const SyrkTile* tiles, int n_tiles, const MvInstance* inst, float* w,
                          const int64_t* w_off, hipStream_t s, bool utu = false);
// S^-1 from the FP64 Cholesky factor of S (w: scratch of the same size).
hipError_t launch_spd_inverse_f64(const double* l, double* w, double* out, const int64_t* off, const int32_t* dims,
                                  int n_inst, hipStream_t s);
// U = T^-1 per (instance, sample, 32-column panel ta); d <= kTrtriMaxDim.
constexpr int kTrtriMaxDim = 512;
hipError_t launch_trtri_panel(const SyrkTile* tiles, int n_tiles, int max_rows, const MvInstance* inst, float* u,
                              const int64_t* u_off, hipStream_t s);
hipError_t launch_dirichlet_norm(const int32_t* dims, const int64_t* samples, const int64_t* g_off,
                                 const float* g, float* out, int n_inst, int64_t max_n, hipStream_t s);

constexpr int kLzVariants = 4;
inline int lz_cols(int variant) { return variant == 1 ? 128 : variant == 2 ? 256 : 64; }
inline int lz_rows(int variant) { return variant == 0 ? 256 : variant == 1 ? 128 : 64; }

}  // namespace erd
