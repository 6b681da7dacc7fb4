// Host orchestration of the batched multivariate generators (C ABI):
// MvNormalGen, WishartGen, DirichletGen over many instances of uneven
// dimension, each instance driven by its own engine.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <new>
#include <mutex>
#include <vector>
#include "eigenrand_amd.h"
#include "erd_capi_internal.hpp"
#include "erd_device.hpp"
#include "erd_kernels.hpp"
#include "erd_mv.hpp"
#include "erd_segment.hpp"
#include "erd_internal.hpp"

namespace erd {

inline int hip_rc(hipError_t e) {
    if (e == hipSuccess) return ERD_OK;
    return e == hipErrorOutOfMemory ? ERD_ERR_OOM : ERD_ERR_HIP;
}

// ------------------------------------------------------------ autotuner ---
// Picks the L*Z tile variant per 64-wide dimension bucket by timing every
// variant on the first batch that contains the bucket; the choice is cached
// for the process. All variants give identical results (same per-element k
// order), so tuning never changes an output. ERD_AUTOTUNE=0 disables it.
std::mutex g_tune_mu;
std::map<int, int> g_tuned;

int default_variant(int dp) { return dp <= 64 ? 2 : dp <= 128 ? 1 : 0; }

bool autotune_enabled() {
    const char* v = std::getenv("ERD_AUTOTUNE");
    return !(v && v[0] == '0');
}

// ERD_LZ_VARIANT=k (0..kLzVariants-1) runs variant k for every bucket,
// bypassing the tuner and its cache (read per call; tests use it to compare
// the variants).
int forced_variant() {
    const char* v = std::getenv("ERD_LZ_VARIANT");
    if (!v || v[0] < '0' || v[0] >= '0' + kLzVariants || v[1] != 0) return -1;
    return v[0] - '0';
}

int bucket_of(int d) { return ((d + 15) & ~15) <= 64 ? 0 : ((d + 15) & ~15) <= 128 ? 1 : ((d + 15) & ~15) <= 256 ? 2 : 3; }

// Tile queues of one L*Z launch. Each instance goes whole to one of the
// kLzQueues queues (largest work first, to the least loaded queue), so the
// workgroups of one XCD read its L planes and Z columns through that XCD's
// L2 only. In a queue: instances largest first (the tail of the launch is
// short), and per instance the row tiles of a column block one after the
// other (they read the same Z columns).
// Bartlett factors (T = L A for the Wishart samples, A lower triangular per
// sample): column blocks stay inside one sample, the product of a block
// whose columns start at i starts at the k chunk holding i (A[k][i] = 0
// for k < i: the skipped terms are exact zeros, the results unchanged), and
// row tiles wholly above the block's first column are not computed. Those
// entries of T are zero and never read: the syrk and the triangular inverse
// read T's upper part only inside 64-aligned diagonal blocks, and every
// skipped tile lies above one (block columns start at multiples of 64).
void build_tiles(const std::vector<MvInstance>& inst, const std::vector<int>& which, int variant,
                 std::vector<MvTile>& tiles, std::vector<int32_t>& qbeg, bool bartlett = false) {
    const int rows = lz_rows(variant), cols = lz_cols(variant);
    std::vector<int> order(which);
    auto work = [&](int i) {
        const int64_t dp = (inst[i].d + 15) & ~15;
        return dp * dp * inst[i].n;
    };
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return work(a) > work(b); });
    std::vector<std::vector<int>> queue(kLzQueues);
    std::vector<int64_t> load(kLzQueues, 0);
    for (int i : order) {
        const int q = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        queue[q].push_back(i);
        load[q] += work(i);
    }
    tiles.clear();
    qbeg.assign(1, 0);
    for (int q = 0; q < kLzQueues; ++q) {
        for (int i : queue[q]) {
            const int d = inst[i].d, dp = (d + 15) & ~15;
            // Per-sample tiles skip a Bartlett factor's zero blocks, which
            // happens only in a sample's second and later column tiles (the
            // first starts at k = 0 and skips no row tile); a sample no wider
            // than one tile takes the tiles across samples instead (the zero
            // blocks are then multiplied, with the same results), so that the
            // tile's columns are used.
            if (!bartlett || d <= cols) {
                for (int64_t c = 0; c < inst[i].n; c += cols)
                    for (int r = 0; r < dp; r += rows)
                        tiles.push_back(MvTile{i, r, c, 0, (int32_t)std::min<int64_t>(cols, inst[i].n - c)});
                continue;
            }
            for (int64_t j0 = 0; j0 < inst[i].n; j0 += d)  // sample j0 / d
                for (int c = 0; c < d; c += cols)
                    for (int r = 0; r < dp; r += rows)
                        if (r + rows > c)
                            tiles.push_back(MvTile{i, r, j0 + c, c & ~31, std::min(cols, d - c)});
        }
        qbeg.push_back((int32_t)tiles.size());
    }
}

int grid_for(int variant, size_t n_tiles) {
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        cus = 256;
    // a persistent grid of the workgroups that fit at once
    return (int)std::min<size_t>(n_tiles, (size_t)cus * lz_blocks_per_cu(variant));
}

// Tile lists of run_lz without launching, for callers that stage their own
// uploads. Returns ERD_ERR_UNSUPPORTED while a dimension bucket is still to
// be tuned (run_lz then tunes it).
int plan_lz(const std::vector<MvInstance>& inst, std::vector<LzLaunch>& plan) {
    std::map<int, std::vector<int>> by_variant;
    for (int i = 0; i < (int)inst.size(); ++i) {
        const int bkt = bucket_of(inst[i].d);
        int v = -1;
        {
            std::lock_guard<std::mutex> lk(g_tune_mu);
            auto it = g_tuned.find(bkt);
            if (it != g_tuned.end()) v = it->second;
        }
        if (forced_variant() >= 0) v = forced_variant();
        if (v < 0) {
            if (autotune_enabled()) return ERD_ERR_UNSUPPORTED;
            v = default_variant(bkt == 0 ? 64 : bkt == 1 ? 128 : 256);
        }
        by_variant[v].push_back(i);
    }
    plan.clear();
    for (auto& kv : by_variant) {
        LzLaunch l;
        l.variant = kv.first;
        build_tiles(inst, kv.second, kv.first, l.tiles, l.qbeg);
        l.grid = grid_for(kv.first, l.tiles.size());
        plan.push_back(std::move(l));
    }
    return ERD_OK;
}

// X = L*Z (+mean) for every instance: group instances by tuned variant, one
// persistent launch per group.
int run_lz(const std::vector<MvInstance>& inst, MvInstance* d_inst, hipStream_t s, Scratch& sc, bool bartlett) {
    std::map<int, std::vector<int>> by_bucket;
    for (int i = 0; i < (int)inst.size(); ++i) by_bucket[bucket_of(inst[i].d)].push_back(i);
    std::map<int, std::vector<int>> by_variant;
    for (auto& kv : by_bucket) {
        int v;
        {
            std::lock_guard<std::mutex> lk(g_tune_mu);
            auto it = g_tuned.find(kv.first);
            v = it != g_tuned.end() ? it->second : -1;
        }
        const int forced = forced_variant();
        if (forced >= 0) v = forced;
        if (v < 0 && autotune_enabled()) {
            float best = 1e30f;
            int best_v = default_variant(kv.first == 0 ? 64 : kv.first == 1 ? 128 : 256);
            for (int cand = 0; cand < kLzVariants; ++cand) {
                std::vector<MvTile> tiles;
                std::vector<int32_t> qbeg;
                build_tiles(inst, kv.second, cand, tiles, qbeg, bartlett);
                MvTile* d_tiles;
                int32_t *d_qbeg, *d_cnt;
                if (sc.upload(&d_tiles, tiles) != hipSuccess || sc.upload(&d_qbeg, qbeg) != hipSuccess ||
                    sc.alloc(&d_cnt, kLzQueues) != hipSuccess)
                    return ERD_ERR_HIP;
                hipEvent_t a = nullptr, b = nullptr;
                float ms = 0;
                hipError_t e = hipEventCreate(&a);
                if (e == hipSuccess) e = hipEventCreate(&b);
                if (e == hipSuccess) e = hipMemsetAsync(d_cnt, 0, sizeof(int32_t) * kLzQueues, s);
                if (e == hipSuccess) e = hipEventRecord(a, s);
                if (e == hipSuccess)
                    e = launch_lz_gemm(cand, grid_for(cand, tiles.size()), d_tiles, d_qbeg, d_cnt, d_inst, s,
                                       bartlett);
                if (e == hipSuccess) e = hipEventRecord(b, s);
                if (e == hipSuccess) e = hipEventSynchronize(b);
                if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
                if (a) (void)hipEventDestroy(a);
                if (b) (void)hipEventDestroy(b);
                if (e != hipSuccess) return ERD_ERR_HIP;
                if (ms < best) {
                    best = ms;
                    best_v = cand;
                }
            }
            v = best_v;
            std::lock_guard<std::mutex> lk(g_tune_mu);
            g_tuned[kv.first] = v;
        }
        if (v < 0) v = default_variant(kv.first == 0 ? 64 : kv.first == 1 ? 128 : 256);
        auto& dst = by_variant[v];
        dst.insert(dst.end(), kv.second.begin(), kv.second.end());
    }
    for (auto& kv : by_variant) {
        std::vector<MvTile> tiles;
        std::vector<int32_t> qbeg;
        build_tiles(inst, kv.second, kv.first, tiles, qbeg, bartlett);
        MvTile* d_tiles;
        int32_t *d_qbeg, *d_cnt;
        if (sc.upload(&d_tiles, tiles) != hipSuccess || sc.upload(&d_qbeg, qbeg) != hipSuccess ||
            sc.alloc(&d_cnt, kLzQueues) != hipSuccess)
            return ERD_ERR_HIP;
        if (hipMemsetAsync(d_cnt, 0, sizeof(int32_t) * kLzQueues, s) != hipSuccess) return ERD_ERR_HIP;
        hipError_t e =
            launch_lz_gemm(kv.first, grid_for(kv.first, tiles.size()), d_tiles, d_qbeg, d_cnt, d_inst, s, bartlett);
        if (e != hipSuccess) return ERD_ERR_HIP;
    }
    return ERD_OK;
}

// FP64 Cholesky of n matrices (device, concatenated column-major, in place),
// then bf16 hi/lo split and float copy of L. Returns ERD_ERR_NOT_PSD if any
// pivot is not positive (info filled per instance).
int factor_and_split(const std::vector<int32_t>& dims, double* d_a, const std::vector<int64_t>& off,
                     uint16_t** lhi, uint16_t** llo, float** l32, std::vector<int64_t>& split_off,
                     std::vector<int32_t>* info_out, hipStream_t s, Scratch& sc) {
    const int n = (int)dims.size();
    int64_t* d_off;
    int32_t *d_dims, *d_info;
    if (sc.upload(&d_off, off) || sc.upload(&d_dims, dims) || sc.alloc(&d_info, n)) return ERD_ERR_HIP;
    int dmax = 0;
    for (int32_t dd : dims) dmax = dd > dmax ? dd : dmax;
    if (launch_cholesky_f64(d_a, d_off, d_dims, d_info, n, dmax, s) != hipSuccess) return ERD_ERR_HIP;
    std::vector<int32_t> info(n);
    if (hipMemcpyAsync(info.data(), d_info, sizeof(int32_t) * n, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return ERD_ERR_HIP;
    if (info_out) *info_out = info;
    for (int v : info)
        if (v) return ERD_ERR_NOT_PSD;
    if (!lhi) return ERD_OK;
    split_off.assign(n + 1, 0);
    int max_dp = 16;
    int64_t total_l = 0;
    for (int i = 0; i < n; ++i) {
        const int dp = (dims[i] + 15) & ~15;
        max_dp = std::max(max_dp, dp);
        split_off[i + 1] = split_off[i] + (int64_t)dp * dp;
        total_l += (int64_t)dims[i] * dims[i];
    }
    int64_t* d_soff;
    if (sc.upload(&d_soff, split_off) || sc.alloc(lhi, split_off[n]) || sc.alloc(llo, split_off[n]) ||
        sc.alloc(l32, total_l))
        return ERD_ERR_HIP;
    if (launch_split_l(d_a, d_off, d_dims, d_soff, *lhi, *llo, *l32, n, max_dp, s) != hipSuccess)
        return ERD_ERR_HIP;
    return ERD_OK;
}

// Stream-ordered release of MvNormal/Wishart factor planes after the work
// on `s`.
int release_factors(erd_mv_factors* f, hipStream_t s) {
    const size_t bytes = sizeof(uint16_t) * (size_t)(f->split_off.empty() ? 0 : f->split_off.back());
    scratch_free(f->lhi, bytes, s);
    scratch_free(f->llo, bytes, s);
    delete f;
    return ERD_OK;
}

// StdNormalGen fills of one engine each, a workgroup per engine (the
// longest first); same outputs and engine states as the segment program.
int run_std_normal_jobs(erd_engines* e, std::vector<FillJob> jobs, hipStream_t s, Scratch& sc) {
    if (jobs.empty()) return ERD_OK;
    std::stable_sort(jobs.begin(), jobs.end(), [](const FillJob& a, const FillJob& b) { return a.count > b.count; });
    FillJob* d_jobs;
    if (sc.upload(&d_jobs, jobs) != hipSuccess) return ERD_ERR_HIP;
    // 640-thread workgroups when all of them are resident at once (three per
    // CU), so the longest fills get ten waves each
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        cus = 256;
    const bool wide = jobs.size() <= (size_t)cus * 3;
    return launch_std_normal_jobs(e->states, d_jobs, (int)jobs.size(), s, wide) == hipSuccess ? ERD_OK : ERD_ERR_HIP;
}

}  // namespace erd

using namespace erd;

extern "C" {

int erd_cholesky_f64_batch(int64_t n_inst, const int32_t* dims, double* a, int32_t* info, void* stream) {
    if (n_inst <= 0 || !dims || !a) return ERD_ERR_INVALID;
    hipStream_t s = (hipStream_t)stream;
    std::vector<int32_t> vd(dims, dims + n_inst);
    std::vector<int64_t> off(n_inst + 1, 0);
    for (int64_t i = 0; i < n_inst; ++i) {
        if (vd[i] <= 0) return ERD_ERR_INVALID;
        off[i + 1] = off[i] + (int64_t)vd[i] * vd[i];
    }
    Scratch sc(s);
    std::vector<int32_t> inf;
    std::vector<int64_t> so;
    int rc = factor_and_split(vd, a, off, nullptr, nullptr, nullptr, so, &inf, s, sc);
    if (info) std::memcpy(info, inf.data(), sizeof(int32_t) * (size_t)n_inst);
    return rc;
}

int erd_mv_factors_create(erd_mv_factors** out, int64_t n_inst, const int32_t* dims, const double* a,
                          int32_t* info, void* stream) {
    if (!out || n_inst <= 0 || !dims || !a) return ERD_ERR_INVALID;
    *out = nullptr;
    hipStream_t s = (hipStream_t)stream;
    std::unique_ptr<erd_mv_factors> f(new (std::nothrow) erd_mv_factors());
    if (!f) return ERD_ERR_OOM;
    f->dims.assign(dims, dims + n_inst);
    std::vector<int64_t> off(n_inst + 1, 0);
    for (int64_t i = 0; i < n_inst; ++i) {
        if (f->dims[i] <= 0) return ERD_ERR_INVALID;
        off[i + 1] = off[i] + (int64_t)f->dims[i] * f->dims[i];
    }
    Scratch sc(s);
    double* d_a;
    if (sc.alloc(&d_a, off[n_inst]) != hipSuccess) return ERD_ERR_OOM;
    if (hipMemcpyAsync(d_a, a, sizeof(double) * off[n_inst], hipMemcpyDeviceToDevice, s) != hipSuccess)
        return ERD_ERR_HIP;
    float* l32;
    std::vector<int32_t> inf;
    int rc = factor_and_split(f->dims, d_a, off, &f->lhi, &f->llo, &l32, f->split_off, &inf, s, sc);
    if (info && !inf.empty()) std::memcpy(info, inf.data(), sizeof(int32_t) * (size_t)n_inst);
    if (rc) return rc;
    sc.keep(f->lhi);
    sc.keep(f->llo);
    *out = f.release();
    return ERD_OK;
}

int erd_mv_factors_create_inverse(erd_mv_factors** out, int64_t n_inst, const int32_t* dims, const double* scale,
                                  int32_t* info, void* stream) {
    if (!out || n_inst <= 0 || !dims || !scale) return ERD_ERR_INVALID;
    *out = nullptr;
    hipStream_t s = (hipStream_t)stream;
    std::vector<int32_t> vd(dims, dims + n_inst);
    std::vector<int64_t> off(n_inst + 1, 0);
    for (int64_t i = 0; i < n_inst; ++i) {
        if (vd[i] <= 0) return ERD_ERR_INVALID;
        off[i + 1] = off[i] + (int64_t)vd[i] * vd[i];
    }
    // Cholesky of the scale (its info is the caller's), S^-1 from it, then the
    // factors of S^-1 as for MvNormalGen.
    Scratch sc(s);
    double *d_a, *d_w, *d_inv;
    int64_t* d_off;
    int32_t* d_dims;
    if (sc.alloc(&d_a, off[n_inst]) || sc.alloc(&d_w, off[n_inst]) || sc.alloc(&d_inv, off[n_inst]) ||
        sc.upload(&d_off, off) || sc.upload(&d_dims, vd))
        return ERD_ERR_OOM;
    if (hipMemcpyAsync(d_a, scale, sizeof(double) * off[n_inst], hipMemcpyDeviceToDevice, s) != hipSuccess)
        return ERD_ERR_HIP;
    std::vector<int32_t> inf;
    std::vector<int64_t> so;
    int rc = factor_and_split(vd, d_a, off, nullptr, nullptr, nullptr, so, &inf, s, sc);
    if (info && !inf.empty()) std::memcpy(info, inf.data(), sizeof(int32_t) * (size_t)n_inst);
    if (rc) return rc;
    if (launch_spd_inverse_f64(d_a, d_w, d_inv, d_off, d_dims, (int)n_inst, s) != hipSuccess) return ERD_ERR_HIP;
    return erd_mv_factors_create(out, n_inst, dims, d_inv, nullptr, stream);
}

int erd_mv_factors_destroy(erd_mv_factors* f) {
    if (!f) return ERD_OK;
    // The planes may still be read by work on any stream: wait for the device.
    if (hipDeviceSynchronize() != hipSuccess) {
        delete f;
        return ERD_ERR_HIP;
    }
    const int rc = release_factors(f, nullptr);
    return rc ? rc : (hipStreamSynchronize(nullptr) == hipSuccess ? ERD_OK : ERD_ERR_HIP);
}

int64_t erd_mv_factors_count(const erd_mv_factors* f) { return f ? (int64_t)f->dims.size() : 0; }

int erd_mvnormal_generate(erd_engines* e, const erd_mv_factors* f, const int64_t* samples, const float* mean,
                          float* out, float* z_out, void* stream) {
    if (layout_unsupported(e)) return ERD_ERR_UNSUPPORTED;
    if (!e || !e->states || !f || !samples || !out) return ERD_ERR_INVALID;
    const int64_t n_inst = (int64_t)f->dims.size();
    if (n_inst > e->count) return ERD_ERR_INVALID;
    hipStream_t s = (hipStream_t)stream;
    const std::vector<int32_t>& vd = f->dims;
    std::vector<int64_t> zoff(n_inst + 1, 0), moff(n_inst + 1, 0);
    for (int64_t i = 0; i < n_inst; ++i) {
        if (samples[i] < 0 || ((int64_t)vd[i] * samples[i]) % 8 != 0) return ERD_ERR_INVALID;
        zoff[i + 1] = zoff[i] + (int64_t)vd[i] * samples[i];
        moff[i + 1] = moff[i] + vd[i];
    }
    Scratch sc(s);
    // Z_i: the StdNormalGen fill of d_i x n_i on engine i (MvNormal.h:103-108).
    float* z = z_out;
    if (!z && sc.alloc(&z, zoff[n_inst]) != hipSuccess) return ERD_ERR_OOM;
    std::vector<FillJob> jobs;
    for (int64_t i = 0; i < n_inst; ++i)
        jobs.push_back(FillJob{(int32_t)i, 0, (int64_t)vd[i] * samples[i], z + zoff[i]});
    int rc = run_std_normal_jobs(e, jobs, s, sc);
    if (rc) return rc;
    std::vector<MvInstance> inst(n_inst);
    for (int64_t i = 0; i < n_inst; ++i)
        inst[i] = MvInstance{vd[i], 0, samples[i], f->lhi + f->split_off[i], f->llo + f->split_off[i],
                             z + zoff[i], out + zoff[i], mean ? mean + moff[i] : nullptr};
    MvInstance* d_inst;
    if (sc.upload(&d_inst, inst) != hipSuccess) return ERD_ERR_HIP;
    return run_lz(inst, d_inst, s, sc);
}

int erd_mvnormal_batch(erd_engines* e, int64_t n_inst, const int32_t* dims, const int64_t* samples,
                       const double* cov, const float* mean, float* out, float* z_out, void* stream) {
    if (layout_unsupported(e)) return ERD_ERR_UNSUPPORTED;
    if (!e || !e->states || n_inst <= 0 || n_inst > e->count || !dims || !samples || !cov || !out)
        return ERD_ERR_INVALID;
    erd_mv_factors* f;
    int rc = erd_mv_factors_create(&f, n_inst, dims, cov, nullptr, stream);
    if (rc) return rc;
    rc = erd_mvnormal_generate(e, f, samples, mean, out, z_out, stream);
    const int rc2 = release_factors(f, (hipStream_t)stream);
    return rc ? rc : rc2;
}

static int wishart_generate(erd_engines* e, const erd_mv_factors* f, const int64_t* df, const int64_t* samples,
                            float* out, void* stream, bool inverse) {
    if (layout_unsupported(e)) return ERD_ERR_UNSUPPORTED;
    if (!e || !e->states || !f || !df || !samples || !out) return ERD_ERR_INVALID;
    const int64_t n_inst = (int64_t)f->dims.size();
    if (n_inst > e->count) return ERD_ERR_INVALID;
    hipStream_t s = (hipStream_t)stream;
    const std::vector<int32_t>& vd = f->dims;
    std::vector<int64_t> aoff(n_inst + 1, 0), nrm_off(n_inst + 1, 0), chi_off(n_inst + 1, 0);
    for (int64_t i = 0; i < n_inst; ++i) {
        const int64_t d = vd[i], n = samples[i];
        // df > d - 1 (MvNormal.h:205); fills need whole packets
        if (n < 0 || df[i] <= d - 1 || n % 8 != 0) return ERD_ERR_INVALID;
        if (inverse && d > kTrtriMaxDim) return ERD_ERR_UNSUPPORTED;
        aoff[i + 1] = aoff[i] + d * d * n;
        nrm_off[i + 1] = nrm_off[i] + n * d * (d - 1) / 2;
        chi_off[i + 1] = chi_off[i] + d * n;
    }
    Scratch sc(s);
    float *nrm, *chi, *amat = nullptr, *tmat;
    if (sc.alloc(&nrm, nrm_off[n_inst]) || sc.alloc(&chi, chi_off[n_inst]) ||
        (inverse && sc.alloc(&amat, aoff[n_inst])) || sc.alloc(&tmat, aoff[n_inst]))
        return ERD_ERR_OOM;
    // Engine program per instance: the normal fill, then chi^2(df - i) for
    // i = 0..d-1 (ChiSquaredGen = GammaGen(n/2, 2)), as generate() draws them.
    // The normal fills run first, one workgroup per engine; the chi^2 rows
    // follow on the same stream, the d fills of an instance in order in one
    // four-wave workgroup (row k: GammaGen((df - k) / 2, 2) over the samples).
    std::vector<FillJob> jobs;
    std::vector<GammaRowGroup> rows((size_t)n_inst);
    std::vector<float> alpha;
    for (int64_t i = 0; i < n_inst; ++i) {
        const int64_t d = vd[i], n = samples[i];
        jobs.push_back(FillJob{(int32_t)i, 0, n * d * (d - 1) / 2, nrm + nrm_off[i]});
        rows[(size_t)i] = GammaRowGroup{(int32_t)i, (int32_t)d, n, chi + chi_off[i], (int64_t)alpha.size(), 2.0f, 0};
        for (int64_t k = 0; k < d; ++k) {
            const float nu = (float)(df[i] - k);
            alpha.push_back(nu * 0.5f);
        }
    }
    std::stable_sort(rows.begin(), rows.end(), [](const GammaRowGroup& x, const GammaRowGroup& y) {
        return (int64_t)x.rows * x.count > (int64_t)y.rows * y.count;
    });
    int rc = run_std_normal_jobs(e, jobs, s, sc);
    if (rc) return rc;
    GammaRowGroup* d_rows;
    float* d_alpha;
    if (sc.upload(&d_rows, rows) || sc.upload(&d_alpha, alpha)) return ERD_ERR_HIP;
    if (launch_gamma_rows(e->states, d_rows, (int32_t)n_inst, d_alpha, guard_of(e), s) != hipSuccess)
        return ERD_ERR_HIP;
    // T = L A with A built from the normals and chi-squares in the product's
    // operand fetch
    std::vector<MvInstance> inst(n_inst);
    for (int64_t i = 0; i < n_inst; ++i)
        inst[i] = MvInstance{vd[i], 0, (int64_t)vd[i] * samples[i], f->lhi + f->split_off[i],
                             f->llo + f->split_off[i], nrm + nrm_off[i], tmat + aoff[i], nullptr, chi + chi_off[i]};
    MvInstance* d_inst;
    int64_t* d_aoff;
    if (sc.upload(&d_inst, inst) || sc.upload(&d_aoff, aoff)) return ERD_ERR_HIP;
    rc = run_lz(inst, d_inst, s, sc, true);
    if (rc) return rc;
    // 64 x 64 output tiles of W (ta >= tb). All tiles of one sample go to
    // one XCD, so that its L2 fetches the sample's T once: workgroups are
    // dispatched to the eight XCDs round-robin, so tile list position p runs
    // on XCD p mod 8. Samples are dealt to the XCDs by least load (work of a
    // tile ~ its k range), and the eight lists are interleaved, padded with
    // empty tiles (inst -1) to equal length.
    std::vector<std::vector<SyrkTile>> bins(kXcds);
    {
        std::vector<double> load(kXcds, 0.0);
        for (int64_t i = 0; i < n_inst; ++i) {
            const int nb = (vd[i] + kSyrkTile - 1) / kSyrkTile;
            for (int64_t j = 0; j < samples[i]; ++j) {
                const int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
                for (int a = 0; a < nb; ++a)
                    for (int b = 0; b <= a; ++b) {
                        bins[x].push_back(SyrkTile{(int32_t)i, (int32_t)j, a, b});
                        load[x] += inverse ? (double)(vd[i] - a * kSyrkTile) : (double)std::min(vd[i], (b + 1) * kSyrkTile);
                    }
            }
        }
    }
    size_t per_bin = 0;
    for (const auto& b : bins) per_bin = std::max(per_bin, b.size());
    std::vector<SyrkTile> tiles(per_bin * kXcds, SyrkTile{-1, 0, 0, 0});
    for (int x = 0; x < kXcds; ++x)
        for (size_t m = 0; m < bins[x].size(); ++m) tiles[m * kXcds + x] = bins[x][m];
    SyrkTile* d_tiles;
    if (sc.upload(&d_tiles, tiles) != hipSuccess) return ERD_ERR_HIP;
    if (inverse) {
        // InvWishartGen (MvNormal.h:427-440): U_j = T_j^-1 into amat, then
        // W_j = U_j^T U_j from U.
        // panel p of an instance with nb blocks solves nb - p block rows: the
        // longest panels go first, and the LDS panel is sized for them
        // (all panels of a sample on one XCD, as the syrk tiles: they all
        // read the sample's T)
        // Each bin lists its panels longest first (nb - p block rows): the
        // samples are dealt to the bins first, then the panels are emitted by
        // length.
        std::vector<std::vector<SyrkTile>> pbins(kXcds);
        std::vector<double> pload(kXcds, 0.0);
        std::vector<int64_t> sbase(n_inst + 1, 0);
        int max_nb = 0;
        for (int64_t i = 0; i < n_inst; ++i) {
            sbase[i + 1] = sbase[i] + samples[i];
            max_nb = std::max(max_nb, (vd[i] + 31) / 32);
        }
        std::vector<uint8_t> xbin((size_t)sbase[n_inst]);
        for (int64_t i = 0; i < n_inst; ++i) {
            const int nb = (vd[i] + 31) / 32;
            for (int64_t j = 0; j < samples[i]; ++j) {
                const int x = (int)(std::min_element(pload.begin(), pload.end()) - pload.begin());
                xbin[(size_t)(sbase[i] + j)] = (uint8_t)x;
                pload[x] += (double)nb * (nb + 1) / 2;
            }
        }
        for (int len = max_nb; len >= 1; --len)
            for (int64_t i = 0; i < n_inst; ++i) {
                const int nb = (vd[i] + 31) / 32;
                if (nb < len) continue;
                for (int64_t j = 0; j < samples[i]; ++j)
                    pbins[xbin[(size_t)(sbase[i] + j)]].push_back(SyrkTile{(int32_t)i, (int32_t)j, nb - len, 0});
            }
        const int max_rows = 32 * max_nb;
        size_t per_pbin = 0;
        for (const auto& b : pbins) per_pbin = std::max(per_pbin, b.size());
        std::vector<SyrkTile> panels(per_pbin * kXcds, SyrkTile{-1, 0, 0, 0});
        for (int x = 0; x < kXcds; ++x)
            for (size_t m = 0; m < pbins[x].size(); ++m) panels[m * kXcds + x] = pbins[x][m];
        std::vector<MvInstance> uinst = inst;
        for (int64_t i = 0; i < n_inst; ++i) uinst[i].x = amat + aoff[i];
        SyrkTile* d_panels;
        MvInstance* d_uinst;
        if (sc.upload(&d_panels, panels) || sc.upload(&d_uinst, uinst)) return ERD_ERR_HIP;
        if (launch_trtri_panel(d_panels, (int)panels.size(), max_rows, d_inst, amat, d_aoff, s) != hipSuccess)
            return ERD_ERR_HIP;
        if (launch_syrk_tt(d_tiles, (int)tiles.size(), d_uinst, out, d_aoff, s, true) != hipSuccess)
            return ERD_ERR_HIP;
        return ERD_OK;
    }
    if (launch_syrk_tt(d_tiles, (int)tiles.size(), d_inst, out, d_aoff, s) != hipSuccess) return ERD_ERR_HIP;
    return ERD_OK;
}

int erd_wishart_generate(erd_engines* e, const erd_mv_factors* f, const int64_t* df, const int64_t* samples,
                         float* out, void* stream) {
    return wishart_generate(e, f, df, samples, out, stream, false);
}

int erd_invwishart_generate(erd_engines* e, const erd_mv_factors* f, const int64_t* df, const int64_t* samples,
                            float* out, void* stream) {
    return wishart_generate(e, f, df, samples, out, stream, true);
}

int erd_invwishart_batch(erd_engines* e, int64_t n_inst, const int32_t* dims, const int64_t* df,
                         const int64_t* samples, const double* scale, float* out, void* stream) {
    if (layout_unsupported(e)) return ERD_ERR_UNSUPPORTED;
    if (!e || !e->states || n_inst <= 0 || n_inst > e->count || !dims || !df || !samples || !scale || !out)
        return ERD_ERR_INVALID;
    for (int64_t i = 0; i < n_inst; ++i) {
        if (dims[i] <= 0 || samples[i] < 0 || df[i] <= dims[i] - 1 || samples[i] % 8 != 0) return ERD_ERR_INVALID;
        if (dims[i] > kTrtriMaxDim) return ERD_ERR_UNSUPPORTED;
    }
    erd_mv_factors* f;
    int rc = erd_mv_factors_create_inverse(&f, n_inst, dims, scale, nullptr, stream);
    if (rc) return rc;
    rc = erd_invwishart_generate(e, f, df, samples, out, stream);
    const int rc2 = release_factors(f, (hipStream_t)stream);
    return rc ? rc : rc2;
}

int erd_wishart_batch(erd_engines* e, int64_t n_inst, const int32_t* dims, const int64_t* df,
                      const int64_t* samples, const double* scale, float* out, void* stream) {
    if (layout_unsupported(e)) return ERD_ERR_UNSUPPORTED;
    if (!e || !e->states || n_inst <= 0 || n_inst > e->count || !dims || !df || !samples || !scale || !out)
        return ERD_ERR_INVALID;
    for (int64_t i = 0; i < n_inst; ++i)
        if (dims[i] <= 0 || samples[i] < 0 || df[i] <= dims[i] - 1 || samples[i] % 8 != 0) return ERD_ERR_INVALID;
    erd_mv_factors* f;
    int rc = erd_mv_factors_create(&f, n_inst, dims, scale, nullptr, stream);
    if (rc) return rc;
    rc = erd_wishart_generate(e, f, df, samples, out, stream);
    const int rc2 = release_factors(f, (hipStream_t)stream);
    return rc ? rc : rc2;
}

int erd_dirichlet_batch(erd_engines* e, int64_t n_inst, const int32_t* dims, const int64_t* samples,
                        const float* alpha, float* out, void* stream) {
    if (layout_unsupported(e)) return ERD_ERR_UNSUPPORTED;
    if (!e || !e->states || n_inst <= 0 || n_inst > e->count || !dims || !samples || !alpha || !out)
        return ERD_ERR_INVALID;
    hipStream_t s = (hipStream_t)stream;
    std::vector<int32_t> vd(dims, dims + n_inst);
    std::vector<int64_t> goff(n_inst + 1, 0), vs(samples, samples + n_inst);
    int64_t aoff = 0;
    for (int64_t i = 0; i < n_inst; ++i) {
        if (vd[i] <= 0 || vs[i] < 0 || vs[i] % 8 != 0) return ERD_ERR_INVALID;
        goff[i + 1] = goff[i] + (int64_t)vd[i] * vs[i];
        aoff += vd[i];
    }
    Scratch sc(s);
    float* g;
    if (sc.alloc(&g, goff[n_inst]) != hipSuccess) return ERD_ERR_OOM;
    // One GammaGen(alpha_k) fill of n per category, in order (Multinomial.h:190-199),
    // instance i on engine i: a workgroup of four waves per instance (the
    // serial chain of a large-K instance runs at four waves' rate), the
    // longest chains first.
    for (int64_t j = 0; j < aoff; ++j)
        if (!(alpha[j] > 0.0f) || !std::isfinite(alpha[j])) return ERD_ERR_INVALID;
    std::vector<GammaRowGroup> rows((size_t)n_inst);
    int64_t a0 = 0;
    for (int64_t i = 0; i < n_inst; ++i) {
        rows[(size_t)i] = GammaRowGroup{(int32_t)i, vd[i], vs[i], g + goff[i], a0, 1.0f, 0};
        a0 += vd[i];
    }
    std::stable_sort(rows.begin(), rows.end(), [](const GammaRowGroup& x, const GammaRowGroup& y) {
        return (int64_t)x.rows * x.count > (int64_t)y.rows * y.count;
    });
    // The longest chains end last: their instances (work >= kLongChain of
    // the longest) run on the caller's stream; the rest run on a second
    // stream at the same time and are normalised there as soon as their rows
    // are done, in the time the long chains still take. Then the long ones
    // are normalised, and the caller's stream waits for the second.
    constexpr double kLongChain = 0.7;
    const double w_max = (double)rows[0].rows * (double)rows[0].count;
    size_t n_long = 0;
    while (n_long < rows.size() && (double)rows[n_long].rows * (double)rows[n_long].count >= kLongChain * w_max)
        ++n_long;
    hipStream_t s2 = nullptr;
    if (n_long == rows.size() || side_stream(&s2) != hipSuccess) n_long = rows.size();
    struct Part {
        std::vector<int32_t> d;
        std::vector<int64_t> n, off;
        int64_t max_n = 0;
        int32_t* d_d = nullptr;
        int64_t *d_n = nullptr, *d_off = nullptr;
    } part[2];
    for (size_t i = 0; i < rows.size(); ++i) {
        Part& q = part[i < n_long ? 0 : 1];
        q.d.push_back(rows[i].rows);
        q.n.push_back(rows[i].count);
        q.off.push_back(rows[i].out - g);
        q.max_n = std::max(q.max_n, rows[i].count);
    }
    GammaRowGroup* d_rows;
    float* d_alpha;
    if (sc.upload(&d_rows, rows) || sc.alloc(&d_alpha, (size_t)aoff) ||
        upload_async(d_alpha, alpha, sizeof(float) * (size_t)aoff, s))
        return ERD_ERR_HIP;
    for (Part& q : part)
        if (!q.d.empty() && (sc.upload(&q.d_d, q.d) || sc.upload(&q.d_n, q.n) || sc.upload(&q.d_off, q.off)))
            return ERD_ERR_HIP;
    // the long chains are enqueued first, so that their workgroups are
    // dispatched ahead of the short ones
    if (n_long < rows.size() && stream_after(s2, s) != hipSuccess) return ERD_ERR_HIP;
    const Part& q = part[0];
    if (launch_gamma_rows(e->states, d_rows, (int32_t)n_long, d_alpha, guard_of(e), s) != hipSuccess ||
        launch_dirichlet_norm(q.d_d, q.d_n, q.d_off, g, out, (int)q.d.size(), q.max_n, s) != hipSuccess)
        return ERD_ERR_HIP;
    if (n_long < rows.size()) {
        // second stream: the shorter chains, then their normalisation
        const Part& r = part[1];
        if (launch_gamma_rows(e->states, d_rows + n_long, (int32_t)(rows.size() - n_long), d_alpha, guard_of(e),
                              s2) != hipSuccess ||
            launch_dirichlet_norm(r.d_d, r.d_n, r.d_off, g, out, (int)r.d.size(), r.max_n, s2) != hipSuccess)
            return ERD_ERR_HIP;
    }
    if (n_long < rows.size() && stream_after(s, s2) != hipSuccess) return ERD_ERR_HIP;
    return ERD_OK;
}

}  // extern "C"

// ------------------------------------------------------------ discrete / multinomial ---
namespace erd {

int build_table(const double* w, int64_t K, HostTable& t) {
    if (K <= 0) return ERD_ERR_INVALID;
    double sum = 0;
    for (int64_t i = 0; i < K; ++i) sum += w[i];
    if (!std::isfinite(sum)) return ERD_ERR_NONFINITE;
    if (K < 16) {
        std::vector<double> c;
        double acc = 0;
        for (int64_t i = 0; i < K; ++i) c.push_back(acc += w[i]);
        for (double p : c) {
            t.cdf.push_back((uint32_t)(p / c.back() * 0x80000000));
            t.fcdf.push_back((float)(p / c.back()));
            t.dcdf.push_back(p / c.back());
        }
        return ERD_OK;
    }
    const size_t msize = (size_t)K;
    const size_t nb = (size_t)std::ceil(std::log2((double)msize));
    const size_t psize = (size_t)1 << nb;
    t.prob.assign(psize, 0);
    t.fprob.assign(psize, 0.0f);
    t.dprob.assign(psize, 0.0);
    t.alias.assign(psize, 0);
    t.bitmask = (uint32_t)(psize - 1);
    const double mean = sum / psize;
    std::vector<double> f(psize, 0.0);
    for (size_t i = 0; i < msize; ++i) f[i] = w[i] / mean;
    const int32_t imax = 0x7FFFFFFF;
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
            for (under = mm; under < psize && f[under] >= 1; ++under) {
            }
            mm = under + 1;
        } else {
            under = over;
        }
        while (over < psize && f[over] < 1) ++over;
    }
    for (; over < psize; ++over)
        if (f[over] >= 1) {
            t.prob[over] = imax;
            t.fprob[over] = 1.0f;
            t.dprob[over] = 1.0;
            t.alias[over] = (int32_t)over;
        }
    if (under < psize) {
        t.prob[under] = imax;
        t.fprob[under] = 1.0f;
        t.dprob[under] = 1.0;
        t.alias[under] = (int32_t)under;
        for (under = mm; under < msize; ++under)
            if (f[under] < 1) {
                t.prob[under] = imax;
                t.fprob[under] = 1.0f;
                t.dprob[under] = 1.0;
                t.alias[under] = (int32_t)under;
            }
    }
    return ERD_OK;
}

// Upload a table and fill the segment's table fields.
int upload_table(const HostTable& t, Segment& sg, Scratch& sc) {
    if (!t.cdf.empty()) {
        uint32_t* d;
        if (sc.upload(&d, t.cdf) != hipSuccess) return ERD_ERR_HIP;
        sg.in0 = d;
        sg.in1 = nullptr;
        sg.d[1] = (double)t.cdf.size();
        sg.d[2] = 0;
    } else {
        int32_t *p, *a;
        if (sc.upload(&p, t.prob) != hipSuccess || sc.upload(&a, t.alias) != hipSuccess) return ERD_ERR_HIP;
        sg.in0 = p;
        sg.in1 = a;
        sg.d[1] = 0;
        sg.d[2] = (double)t.bitmask;
    }
    return ERD_OK;
}

std::vector<float> multinomial_row_p(const double* w, int64_t K) {
    std::vector<double> probs(w, w + K);
    double sum = 0;
    for (double v : probs) sum += v;
    for (double& v : probs) v /= sum;
    std::vector<float> p;
    double rest_p = 1 - probs[0];
    for (int64_t row = 0; row + 1 < K; ++row) {
        p.push_back(row == 0 ? (float)probs[0] : (float)(probs[row] / rest_p));
        if (row > 0) rest_p -= probs[row];
    }
    return p;
}

}  // namespace erd

extern "C" {

int erd_alias_table(const double* weights, int64_t K, int32_t* prob, int32_t* alias, uint32_t* cdf) {
    if (!weights || K <= 0) return ERD_ERR_INVALID;
    HostTable t;
    int rc = build_table(weights, K, t);
    if (rc) return rc;
    if (!t.cdf.empty()) {
        if (cdf) std::memcpy(cdf, t.cdf.data(), 4 * t.cdf.size());
    } else {
        if (prob) std::memcpy(prob, t.prob.data(), 4 * t.prob.size());
        if (alias) std::memcpy(alias, t.alias.data(), 4 * t.alias.size());
    }
    return ERD_OK;
}

int erd_fill_discrete_i32(erd_engines* e, int32_t* out, int64_t per, const double* weights, int64_t K,
                          void* stream) {
    if (layout_unsupported(e)) return ERD_ERR_UNSUPPORTED;
    if (!e || !e->states || per < 0 || per % 4 != 0 || (per > 0 && !out) || !weights) return ERD_ERR_INVALID;
    hipStream_t s = (hipStream_t)stream;
    HostTable t;
    int rc = build_table(weights, K, t);
    if (rc) return rc;
    Scratch sc(s);
    uint32_t* cdf = nullptr;
    int32_t *prob = nullptr, *alias = nullptr;
    if (!t.cdf.empty()) {
        if (sc.upload(&cdf, t.cdf) != hipSuccess) return ERD_ERR_HIP;
    } else if (sc.upload(&prob, t.prob) != hipSuccess || sc.upload(&alias, t.alias) != hipSuccess) {
        return ERD_ERR_HIP;
    }
    return launch_fill_discrete_i32(e->states, e->count, out, per, cdf, prob, alias, (int32_t)t.cdf.size(),
                                    t.bitmask, s) == hipSuccess
               ? ERD_OK
               : ERD_ERR_HIP;
}

int erd_fill_discrete_f_i32(erd_engines* e, int32_t* out, int64_t per, const double* weights, int64_t K,
                            void* stream) {
    if (layout_unsupported(e)) return ERD_ERR_UNSUPPORTED;
    if (!e || !e->states || per < 0 || per % 4 != 0 || (per > 0 && !out) || !weights) return ERD_ERR_INVALID;
    hipStream_t s = (hipStream_t)stream;
    HostTable t;
    int rc = build_table(weights, K, t);
    if (rc) return rc;
    Scratch sc(s);
    float* tab;
    int32_t* alias = nullptr;
    if (!t.fcdf.empty()) {
        if (sc.upload(&tab, t.fcdf) != hipSuccess) return ERD_ERR_HIP;
    } else if (sc.upload(&tab, t.fprob) != hipSuccess || sc.upload(&alias, t.alias) != hipSuccess) {
        return ERD_ERR_HIP;
    }
    return launch_fill_discrete_f_i32(e->states, e->count, out, per, tab, alias, (int32_t)t.fcdf.size(), t.bitmask,
                                      s) == hipSuccess
               ? ERD_OK
               : ERD_ERR_HIP;
}

int erd_fill_discrete_d_i32(erd_engines* e, int32_t* out, int64_t per, const double* weights, int64_t K,
                            void* stream) {
    if (layout_unsupported(e)) return ERD_ERR_UNSUPPORTED;
    if (!e || !e->states || per < 0 || per % 4 != 0 || (per > 0 && !out) || !weights) return ERD_ERR_INVALID;
    hipStream_t s = (hipStream_t)stream;
    HostTable t;
    int rc = build_table(weights, K, t);
    if (rc) return rc;
    Scratch sc(s);
    double *cdf = nullptr, *prob = nullptr;
    int32_t* alias = nullptr;
    if (!t.dcdf.empty()) {
        if (sc.upload(&cdf, t.dcdf) != hipSuccess) return ERD_ERR_HIP;
    } else if (sc.upload(&prob, t.dprob) != hipSuccess || sc.upload(&alias, t.alias) != hipSuccess) {
        return ERD_ERR_HIP;
    }
    return launch_fill_discrete_d_i32(e->states, e->count, out, per, cdf, (int32_t)t.dcdf.size(), prob, alias,
                                      t.bitmask, s) == hipSuccess
               ? ERD_OK
               : ERD_ERR_HIP;
}

int erd_multinomial_batch(erd_engines* e, int64_t n_inst, const int64_t* dims, const int32_t* trials,
                          const int64_t* samples, const double* weights, int32_t* out, void* stream) {
    if (layout_unsupported(e)) return ERD_ERR_UNSUPPORTED;
    if (!e || !e->states || n_inst <= 0 || n_inst > e->count || !dims || !trials || !samples || !weights ||
        !out)
        return ERD_ERR_INVALID;
    hipStream_t s = (hipStream_t)stream;
    Scratch sc(s);
    std::vector<Segment> segs;
    std::vector<int32_t> eng;
    int64_t woff = 0, ooff = 0;
    for (int64_t i = 0; i < n_inst; ++i) {
        const int64_t K = dims[i], n = samples[i];
        const int32_t T = trials[i];
        if (K <= 0 || n < 0 || T < 0) return ERD_ERR_INVALID;
        for (int64_t k = 0; k < K; ++k)
            if (weights[woff + k] < 0) return ERD_ERR_INVALID;  // eigen_assert(probs[i] >= 0)
        HostTable t;
        int rc = build_table(weights + woff, K, t);
        if (rc) return rc;
        int32_t* o = out + ooff;
        if (T < std::max<int64_t>(30 * (K - 1), 100)) {
            if (hipMemsetAsync(o, 0, sizeof(int32_t) * K * n, s) != hipSuccess) return ERD_ERR_HIP;
            Segment sg{};
            sg.kind = SEG_MULTINOMIAL_SMALL;
            sg.count = n;
            sg.out = o;
            sg.i0 = T;
            sg.d[0] = (double)K;
            rc = upload_table(t, sg, sc);
            if (rc) return rc;
            segs.push_back(sg);
            eng.push_back((int32_t)i);
        } else {
            if (n % 8 != 0) return ERD_ERR_INVALID;  // BinomialVGen<float> rows run in Packet8f
            int32_t* r;  // remaining trials per sample, written by the first row
            if (sc.alloc(&r, (size_t)n) != hipSuccess) return ERD_ERR_OOM;
            float* prow;
            if (sc.upload(&prow, multinomial_row_p(weights + woff, K)) != hipSuccess) return ERD_ERR_HIP;
            Segment sg{};
            sg.kind = SEG_MULTINOMIAL_ROWS;
            sg.count = n;
            sg.out = o;
            sg.in0 = r;
            sg.in1 = prow;
            sg.i0 = T;
            sg.d[1] = (double)K;
            segs.push_back(sg);
            eng.push_back((int32_t)i);
        }
        woff += K;
        ooff += K * n;
    }
    return run_segments(e, segs, eng, s);
}

}  // extern "C"
