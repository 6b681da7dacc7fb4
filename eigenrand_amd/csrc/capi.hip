// The C ABI declared in include/eigenrand_amd.h: argument checks, engine-set
// lifetime, and dispatch to the kernel launchers.
#include <new>
#include <vector>
#include <cmath>
#include <cstring>
#include <algorithm>
#include "eigenrand_amd.h"
#include "erd_device.hpp"
#include "erd_kernels.hpp"
#include "erd_capi_internal.hpp"
#include "erd_internal.hpp"
#include "erd_segment.hpp"

using erd::EngineState;

namespace {

inline hipStream_t S(void* s) { return reinterpret_cast<hipStream_t>(s); }

inline int hip_status(hipError_t e) {
    if (e == hipSuccess) return ERD_OK;
    if (e == hipErrorOutOfMemory) return ERD_ERR_OOM;
    return ERD_ERR_HIP;
}

inline bool valid_set(const erd_engines* e) { return e && e->states && e->count > 0; }

// Common checks of a fill call: a valid set, a multiple of the packet width,
// and a device output pointer whenever something is written.
inline int check_fill(const erd_engines* e, const void* out, int64_t per, int64_t width) {
    if (!valid_set(e)) return ERD_ERR_INVALID;
    if (e->layout != 8) return ERD_ERR_UNSUPPORTED;
    if (per < 0 || per % width != 0) return ERD_ERR_INVALID;
    if (per > 0 && !out) return ERD_ERR_INVALID;
    return ERD_OK;
}

// The same for a fill with a form at every lane layout: its packet width at
// layouts 4, 8 and 16.
inline int check_fill_l(const erd_engines* e, const void* out, int64_t per, int64_t w4, int64_t w8, int64_t w16) {
    if (!valid_set(e)) return ERD_ERR_INVALID;
    const int64_t width = e->layout == 4 ? w4 : e->layout == 16 ? w16 : w8;
    if (per < 0 || per % width != 0) return ERD_ERR_INVALID;
    if (per > 0 && !out) return ERD_ERR_INVALID;
    return ERD_OK;
}

// A fill at layout 4, or at layout 16 where its element order differs.
inline int layout_fill(erd_engines* e, int kind, void* out, int64_t per, double a, double b, int mode,
                       void* stream) {
    return hip_status(erd::launch_fill_layout(e->states, e->count, e->layout, kind, out, per, a, b, mode, S(stream)));
}

}  // namespace

extern "C" {

const char* erd_status_string(int status) {
    switch (status) {
        case ERD_OK: return "ok";
        case ERD_ERR_INVALID: return "invalid argument";
        case ERD_ERR_NONFINITE: return "non-finite weights";
        case ERD_ERR_NOT_PSD: return "matrix is not positive semi-definite";
        case ERD_ERR_ITER_CAP: return "rejection loop hit the iteration cap";
        case ERD_ERR_OOM: return "device out of memory";
        case ERD_ERR_HIP: return "HIP runtime error";
        case ERD_ERR_UNSUPPORTED: return "unsupported parameters";
        default: return "unknown status";
    }
}

size_t erd_engine_state_bytes(void) { return sizeof(EngineState); }

int erd_engines_create(erd_engines** out, const uint64_t* seeds, int64_t n, void* stream) {
    if (!out || !seeds || n <= 0) return ERD_ERR_INVALID;
    *out = nullptr;
    erd_engines* e = new (std::nothrow) erd_engines();
    if (!e) return ERD_ERR_OOM;
    uint64_t* d_seeds = nullptr;
    hipError_t err = hipMalloc(&e->states, sizeof(EngineState) * (size_t)n);
    if (err == hipSuccess) err = hipMalloc(&e->status, sizeof(uint32_t));
    if (err == hipSuccess) err = hipMalloc(&d_seeds, sizeof(uint64_t) * (size_t)n);
    if (err == hipSuccess) err = hipMemsetAsync(e->status, 0, sizeof(uint32_t), S(stream));
    if (err == hipSuccess)
        err = hipMemcpyAsync(d_seeds, seeds, sizeof(uint64_t) * (size_t)n, hipMemcpyHostToDevice, S(stream));
    if (err == hipSuccess) err = erd::launch_seed(e->states, d_seeds, n, S(stream));
    if (err == hipSuccess) err = hipStreamSynchronize(S(stream));
    // the tangent samplers' plgamma table of this device, built here (outside
    // any stream capture) rather than on the first fill
    if (err == hipSuccess && !erd::lgamma_table(S(stream))) err = hipErrorOutOfMemory;
    if (d_seeds) (void)hipFree(d_seeds);
    if (err != hipSuccess) {
        if (e->states) (void)hipFree(e->states);
        if (e->status) (void)hipFree(e->status);
        delete e;
        return hip_status(err);
    }
    e->count = n;
    e->owns = true;
    *out = e;
    return ERD_OK;
}

int erd_engines_create_range(erd_engines** out, uint64_t seed0, uint64_t stride, int64_t n,
                             void* stream) {
    if (!out || n <= 0) return ERD_ERR_INVALID;
    std::vector<uint64_t> seeds((size_t)n);
    for (int64_t g = 0; g < n; ++g) seeds[(size_t)g] = seed0 + (uint64_t)g * stride;
    return erd_engines_create(out, seeds.data(), n, stream);
}

int erd_engines_view(erd_engines** out, erd_engines* parent, int64_t first, int64_t n) {
    if (!out || !valid_set(parent) || first < 0 || n <= 0 || first + n > parent->count)
        return ERD_ERR_INVALID;
    erd_engines* e = new (std::nothrow) erd_engines();
    if (!e) return ERD_ERR_OOM;
    e->states = parent->states + first;
    e->status = parent->status;
    e->reject_cap = parent->reject_cap;
    e->lockstep_cap = parent->lockstep_cap;
    e->layout = parent->layout;
    e->count = n;
    e->owns = false;
    *out = e;
    return ERD_OK;
}

int erd_engines_destroy(erd_engines* e) {
    if (!e) return ERD_OK;
    hipError_t err = hipSuccess;
    if (e->owns) {
        err = hipFree(e->states);
        hipError_t e2 = hipFree(e->status);
        if (err == hipSuccess) err = e2;
    }
    delete e;
    return hip_status(err);
}

int64_t erd_engines_count(const erd_engines* e) { return e ? e->count : 0; }

void* erd_engines_device_ptr(erd_engines* e) { return e ? (void*)e->states : nullptr; }

int erd_engines_snapshot(const erd_engines* e, void* host_dst, void* stream) {
    if (!valid_set(e) || !host_dst) return ERD_ERR_INVALID;
    hipError_t err = hipMemcpyAsync(host_dst, e->states, sizeof(EngineState) * (size_t)e->count,
                                    hipMemcpyDeviceToHost, S(stream));
    if (err == hipSuccess) err = hipStreamSynchronize(S(stream));
    return hip_status(err);
}

int erd_engines_restore(erd_engines* e, const void* host_src, void* stream) {
    if (!valid_set(e) || !host_src) return ERD_ERR_INVALID;
    // Reject records whose block position or cache flag is out of range: a
    // kernel would index LDS with them.
    const EngineState* rec = static_cast<const EngineState*>(host_src);
    for (int64_t g = 0; g < e->count; ++g) {
        if (rec[g].idx < 0 || rec[g].idx > erd::kN) return ERD_ERR_INVALID;
        if (rec[g].half_valid != 0 && rec[g].half_valid != 1) return ERD_ERR_INVALID;
    }
    hipError_t err = hipMemcpyAsync(e->states, host_src, sizeof(EngineState) * (size_t)e->count,
                                    hipMemcpyHostToDevice, S(stream));
    if (err == hipSuccess) err = hipStreamSynchronize(S(stream));
    return hip_status(err);
}

int erd_engines_copy(erd_engines* dst, const erd_engines* src, void* stream) {
    if (!valid_set(dst) || !valid_set(src) || dst->count != src->count) return ERD_ERR_INVALID;
    return hip_status(hipMemcpyAsync(dst->states, src->states, sizeof(EngineState) * (size_t)src->count,
                                     hipMemcpyDeviceToDevice, S(stream)));
}

int erd_engines_discard(erd_engines* e, int64_t calls, void* stream) {
    if (!valid_set(e) || calls < 0) return ERD_ERR_INVALID;
    return hip_status(erd::launch_discard(e->states, e->count, calls, S(stream)));
}

int erd_engines_set_iteration_caps(erd_engines* e, int32_t reject_cap, int32_t lockstep_cap) {
    if (!valid_set(e) || reject_cap < 0 || lockstep_cap < 0) return ERD_ERR_INVALID;
    e->reject_cap = reject_cap > 0 ? reject_cap : erd::kDefaultRejectCap;
    e->lockstep_cap = lockstep_cap > 0 ? lockstep_cap : erd::kDefaultLockstepCap;
    return ERD_OK;
}

int erd_engines_status(erd_engines* e, uint32_t* status_out, void* stream) {
    if (!valid_set(e) || !status_out) return ERD_ERR_INVALID;
    uint32_t v = 0;
    hipError_t err = hipMemcpyAsync(&v, e->status, sizeof(uint32_t), hipMemcpyDeviceToHost, S(stream));
    if (err == hipSuccess) err = hipStreamSynchronize(S(stream));
    if (err == hipSuccess) err = hipMemsetAsync(e->status, 0, sizeof(uint32_t), S(stream));
    if (err == hipSuccess) err = hipStreamSynchronize(S(stream));
    *status_out = v;
    return hip_status(err);
}

int erd_engines_set_layout(erd_engines* e, int layout) {
    if (!valid_set(e) || (layout != 4 && layout != 8 && layout != 16)) return ERD_ERR_INVALID;
    e->layout = layout;
    return ERD_OK;
}

int erd_engines_layout(const erd_engines* e) { return valid_set(e) ? e->layout : 0; }

int erd_fill_u64(erd_engines* e, uint64_t* out, int64_t per, void* stream) {
    int rc = check_fill_l(e, out, per, 2, 4, 4);
    if (rc) return rc;
    if (e->layout == 4) return layout_fill(e, erd::LK_U64, out, per, 0, 0, 0, stream);
    return hip_status(erd::launch_fill_u64(e->states, e->count, out, per, S(stream)));
}

int erd_engines_create_p8(erd_engines** out, const uint64_t* seeds, int64_t n_p8, void* stream) {
    if (!out || !seeds || n_p8 <= 0) return ERD_ERR_INVALID;
    std::vector<uint64_t> vs((size_t)(2 * n_p8));
    for (int64_t k = 0; k < n_p8; ++k) {
        vs[2 * k] = seeds[k];           // rngs[0]{seed}
        vs[2 * k + 1] = seeds[k] + 4u;  // rngs[1]{seed + 1 * u64_stride}, u64_stride = 4
    }
    return erd_engines_create(out, vs.data(), 2 * n_p8, stream);
}

static int p8_fill(erd_engines* e, void* out, int64_t per, int64_t width, int mode, void* stream) {
    int rc = check_fill(e, out, per, width);
    if (rc) return rc;
    if (e->count % 2 != 0) return ERD_ERR_INVALID;  // adaptors are engine pairs
    return hip_status(erd::launch_fill_p8(e->states, e->count, out, per / 2, mode, S(stream)));
}

// One segment per adaptor on the P8 program kernel: the adaptor's scalar
// stream packed as the distribution's packet code takes it (mode -1), or raw
// draws (mode 0).
static int p8_dist(erd_engines* e, erd_request r, int64_t per, int mode, void* stream) {
    if (!e || !e->states || e->count < 2 || e->count % 2 != 0 || per < 0 || (per > 0 && !r.out))
        return ERD_ERR_INVALID;
    if (per == 0) return ERD_OK;
    const int64_t n_p8 = e->count / 2;
    int elem = 4;
    switch (r.dist) {
        case ERD_DIST_U64: elem = 8; break;
        default: break;
    }
    std::vector<erd::Segment> segs((size_t)n_p8);
    std::vector<int32_t> eng((size_t)n_p8);
    for (int64_t k = 0; k < n_p8; ++k) {
        erd_request q = r;
        q.engine = (int32_t)k;
        q.count = per;
        q.out = (char*)r.out + (size_t)k * (size_t)per * (size_t)elem;
        int rc = erd::make_segment(q, segs[(size_t)k]);
        if (rc) return rc;
        if (mode < 0 && erd::p8_feed_mode(segs[(size_t)k].kind) < 0) return ERD_ERR_UNSUPPORTED;
        eng[(size_t)k] = (int32_t)k;
    }
    return erd::run_segments(e, segs, eng, S(stream), true, mode);
}

int erd_p8_fill_dist(erd_engines* e, int32_t dist, void* out, int64_t per_p8, double a, double b, void* stream) {
    erd_request r{};
    r.dist = dist;
    r.out = out;
    r.a = a;
    r.b = b;
    if (dist == ERD_DIST_U64) return ERD_ERR_UNSUPPORTED;  // erd_p8_fill_u64
    return p8_dist(e, r, per_p8, -1, stream);
}

int erd_p8_fill_u64(erd_engines* e, uint64_t* out, int64_t per, void* stream) {
    // the scalar stream itself, from the operator() buffer's position
    erd_request r{};
    r.dist = ERD_DIST_U64;
    r.out = out;
    return p8_dist(e, r, per, 0, stream);
}

int erd_p8_fill_u32(erd_engines* e, uint32_t* out, int64_t per, void* stream) {
    return p8_fill(e, out, per, 16, 1, stream);
}

int erd_p8_fill_uniform_f32(erd_engines* e, float* out, int64_t per, void* stream) {
    return p8_fill(e, out, per, 16, 2, stream);
}

int erd_fill_uniform_f32(erd_engines* e, float* out, int64_t per, float a, float b, void* stream) {
    int rc = check_fill_l(e, out, per, 4, 8, 16);
    if (rc) return rc;
    const int affine = !(a == 0.0f && b == 1.0f);
    if (e->layout == 4) return layout_fill(e, erd::LK_UNIFORM_F32, out, per, a, b, affine, stream);
    return hip_status(erd::launch_fill_uniform_f32(e->states, e->count, out, per, a, b, affine, S(stream)));
}

int erd_fill_uniform_f64(erd_engines* e, double* out, int64_t per, double a, double b, void* stream) {
    int rc = check_fill_l(e, out, per, 2, 4, 8);
    if (rc) return rc;
    const int affine = !(a == 0.0 && b == 1.0);
    if (e->layout == 4) return layout_fill(e, erd::LK_UNIFORM_F64, out, per, a, b, affine, stream);
    return hip_status(erd::launch_fill_uniform_f64(e->states, e->count, out, per, a, b, affine, S(stream)));
}

int erd_fill_std_normal_f64(erd_engines* e, double* out, int64_t per, void* stream) {
    int rc = check_fill_l(e, out, per, 2, 4, 8);
    if (rc) return rc;
    if (e->layout != 8) return layout_fill(e, erd::LK_NORMAL_F64, out, per, 0.0, 1.0, 0, stream);
    return hip_status(erd::launch_fill_normal_f64(e->states, e->count, out, per, 0.0, 1.0, 0, S(stream)));
}

int erd_fill_normal_f64(erd_engines* e, double* out, int64_t per, double mean, double sd, void* stream) {
    int rc = check_fill_l(e, out, per, 2, 4, 8);
    if (rc) return rc;
    if (e->layout != 8) return layout_fill(e, erd::LK_NORMAL_F64, out, per, mean, sd, 1, stream);
    return hip_status(erd::launch_fill_normal_f64(e->states, e->count, out, per, mean, sd, 1, S(stream)));
}

int erd_fill_exponential_f64(erd_engines* e, double* out, int64_t per, double lambda, void* stream) {
    int rc = check_fill_l(e, out, per, 2, 4, 8);
    if (rc) return rc;
    if (e->layout == 4) return layout_fill(e, erd::LK_EXPONENTIAL_F64, out, per, lambda, 0, 0, stream);
    return hip_status(erd::launch_fill_exponential_f64(e->states, e->count, out, per, lambda, S(stream)));
}

int erd_fill_cauchy_f64(erd_engines* e, double* out, int64_t per, double a, double b, void* stream) {
    int rc = check_fill_l(e, out, per, 2, 4, 8);
    if (rc) return rc;
    if (e->layout == 4) return layout_fill(e, erd::LK_CAUCHY_F64, out, per, a, b, 0, stream);
    return hip_status(erd::launch_fill_cauchy_f64(e->states, e->count, out, per, a, b, S(stream)));
}

int erd_fill_std_normal_f32(erd_engines* e, float* out, int64_t per, void* stream) {
    int rc = check_fill_l(e, out, per, 4, 8, 16);
    if (rc) return rc;
    if (e->layout != 8) return layout_fill(e, erd::LK_NORMAL_F32, out, per, 0.0, 1.0, 0, stream);
    return hip_status(erd::launch_fill_normal_f32(e->states, e->count, out, per, 0.f, 1.f, 0, S(stream)));
}

int erd_fill_normal_f32(erd_engines* e, float* out, int64_t per, float mean, float sd, void* stream) {
    int rc = check_fill_l(e, out, per, 4, 8, 16);
    if (rc) return rc;
    if (e->layout != 8) return layout_fill(e, erd::LK_NORMAL_F32, out, per, mean, sd, 1, stream);
    return hip_status(erd::launch_fill_normal_f32(e->states, e->count, out, per, mean, sd, 1, S(stream)));
}

int erd_fill_std_normal_f32_fast(erd_engines* e, float* out, int64_t per, void* stream) {
    int rc = check_fill(e, out, per, 8);
    if (rc) return rc;
    return hip_status(erd::launch_fill_normal_f32_fast(e->states, e->count, out, per, 0.f, 1.f, 0, S(stream)));
}

int erd_fill_normal_f32_fast(erd_engines* e, float* out, int64_t per, float mean, float sd, void* stream) {
    int rc = check_fill(e, out, per, 8);
    if (rc) return rc;
    return hip_status(erd::launch_fill_normal_f32_fast(e->states, e->count, out, per, mean, sd, 1, S(stream)));
}

int erd_store_ceiling_f32(float* out, int64_t n_blocks, int64_t per_block, void* stream) {
    if (n_blocks < 0 || per_block < 0 || per_block % 4 != 0 || n_blocks > 0x7fffffff) return ERD_ERR_INVALID;
    if (n_blocks > 0 && per_block > 0 && !out) return ERD_ERR_INVALID;
    return hip_status(erd::launch_store_blocks(out, n_blocks, per_block, S(stream)));
}

int erd_fill_normal_vop_f32(erd_engines* e, float* out, int64_t per, const float* mean,
                            const float* sd, void* stream) {
    int rc = check_fill(e, out, per, 8);
    if (rc) return rc;
    if (per > 0 && (!mean || !sd)) return ERD_ERR_INVALID;
    return hip_status(erd::launch_fill_normal_vop_f32(e->states, e->count, out, per, mean, sd, S(stream)));
}

int erd_fill_normal_vop_vs_f32(erd_engines* e, float* out, int64_t per, const float* mean, float sd,
                               void* stream) {
    int rc = check_fill(e, out, per, 8);
    if (rc) return rc;
    if (per > 0 && !mean) return ERD_ERR_INVALID;
    return hip_status(erd::launch_fill_normal_vop_f32(e->states, e->count, out, per, mean, nullptr, S(stream),
                                                      0.f, sd));
}

int erd_fill_normal_vop_sv_f32(erd_engines* e, float* out, int64_t per, float mean, const float* sd,
                               void* stream) {
    int rc = check_fill(e, out, per, 8);
    if (rc) return rc;
    if (per > 0 && !sd) return ERD_ERR_INVALID;
    return hip_status(erd::launch_fill_normal_vop_f32(e->states, e->count, out, per, nullptr, sd, S(stream),
                                                      mean, 1.f));
}

int erd_fill_exponential_f32(erd_engines* e, float* out, int64_t per, float lambda, void* stream) {
    int rc = check_fill_l(e, out, per, 4, 8, 16);
    if (rc) return rc;
    if (e->layout == 4) return layout_fill(e, erd::LK_EXPONENTIAL_F32, out, per, lambda, 0, 0, stream);
    return hip_status(erd::launch_fill_exponential_f32(e->states, e->count, out, per, lambda, S(stream)));
}

int erd_fill_lognormal_f32(erd_engines* e, float* out, int64_t per, float mean, float sd, void* stream) {
    int rc = check_fill_l(e, out, per, 4, 8, 16);
    if (rc) return rc;
    if (e->layout != 8) return layout_fill(e, erd::LK_NORMAL_F32, out, per, mean, sd, 3, stream);
    return hip_status(erd::launch_fill_lognormal_f32(e->states, e->count, out, per, mean, sd, S(stream)));
}

int erd_fill_weibull_f32(erd_engines* e, float* out, int64_t per, float a, float b, void* stream) {
    int rc = check_fill_l(e, out, per, 4, 8, 16);
    if (rc) return rc;
    if (e->layout == 4) return layout_fill(e, erd::LK_WEIBULL_F32, out, per, a, b, 0, stream);
    return hip_status(erd::launch_fill_weibull_f32(e->states, e->count, out, per, a, b, S(stream)));
}

int erd_fill_extreme_value_f32(erd_engines* e, float* out, int64_t per, float a, float b, void* stream) {
    int rc = check_fill_l(e, out, per, 4, 8, 16);
    if (rc) return rc;
    if (e->layout == 4) return layout_fill(e, erd::LK_EXTREME_VALUE_F32, out, per, a, b, 0, stream);
    return hip_status(erd::launch_fill_extreme_value_f32(e->states, e->count, out, per, a, b, S(stream)));
}

int erd_fill_cauchy_f32(erd_engines* e, float* out, int64_t per, float a, float b, void* stream) {
    int rc = check_fill_l(e, out, per, 4, 8, 16);
    if (rc) return rc;
    if (e->layout == 4) return layout_fill(e, erd::LK_CAUCHY_F32, out, per, a, b, 0, stream);
    return hip_status(erd::launch_fill_cauchy_f32(e->states, e->count, out, per, a, b, S(stream)));
}

int erd_fill_studentt_f32(erd_engines* e, float* out, int64_t per, float n, void* stream) {
    int rc = check_fill_l(e, out, per, 4, 8, 16);
    if (rc) return rc;
    if (e->layout != 8) return layout_fill(e, erd::LK_STUDENT_T_F32, out, per, n, 0, 0, stream);
    return hip_status(erd::launch_fill_studentt_f32(e->states, e->count, out, per, n, S(stream)));
}

}  // extern "C"

namespace {

// The fills of GammaGen, PoissonGen and BinomialGen; gs (optional) holds the
// CompressMask queues of a reused generator object (rejection regimes only:
// the other regimes have no state).
// Up to this many engines, a rejection-regime gamma fill serves each engine
// with a workgroup of four waves (op_reject_wg) instead of one wave: with
// fewer waves than SIMDs to spread them over, each engine's serial chain is
// the bound.
constexpr int64_t kGammaRowsMaxEngines = 1024;

int fill_gamma(erd_engines* e, float* out, int64_t per, float alpha, float beta, erd::GenState* gs,
               void* stream) {
    int rc = check_fill(e, out, per, 8);
    if (rc) return rc;
    if (!(alpha > 0.0f) || !std::isfinite(alpha)) return ERD_ERR_INVALID;
    if (alpha == 1.0f || ((float)(int)alpha == alpha && alpha < 20.0f))
        return hip_status(erd::launch_fill_gamma_fixed_f32(e->states, e->count, out, per, alpha, beta, S(stream)));
    if (!gs && e->count <= kGammaRowsMaxEngines) {
        // few engines: a four-wave workgroup per engine runs each serial fill
        std::vector<erd::GammaRowGroup> rows((size_t)e->count);
        for (int64_t g = 0; g < e->count; ++g)
            rows[(size_t)g] = erd::GammaRowGroup{(int32_t)g, 1, per, out + g * per, 0, beta, 0};
        erd::Scratch sc(S(stream));
        erd::GammaRowGroup* d_rows;
        float* d_alpha;
        if (sc.upload(&d_rows, rows) || sc.alloc(&d_alpha, 1) ||
            erd::upload_async(d_alpha, &alpha, sizeof(float), S(stream)))
            return ERD_ERR_HIP;
        return hip_status(erd::launch_gamma_rows(e->states, d_rows, (int32_t)e->count, d_alpha, guard_of(e),
                                                 S(stream)));
    }
    return hip_status(erd::launch_fill_gamma_reject_f32(e->states, e->count, out, per, alpha, beta,
                                                        guard_of(e), S(stream), gs));
}

int fill_poisson(erd_engines* e, int32_t* out, int64_t per, double mean, erd::GenState* gs, void* stream) {
    int rc = check_fill(e, out, per, 4);
    if (rc) return rc;
    if (!(mean >= 0.0) || !std::isfinite(mean)) return ERD_ERR_INVALID;
    if (mean < 12) {
        return hip_status(erd::launch_fill_poisson_knuth_i32(e->states, e->count, out, per, mean, 0, 0,
                                                             guard_of(e), S(stream)));
    }
    // PoissonGen constructor (Discrete.h:807-813), in double with libm.
    erd::PoissonParams pp;
    pp.mean = mean;
    pp.ne_mean = std::exp(-mean);
    pp.sqrt_tmean = std::sqrt(2 * mean);
    pp.log_mean = std::log(mean);
    pp.g1 = mean * pp.log_mean - std::lgamma(mean + 1);
    return hip_status(erd::launch_fill_poisson_reject_i32(e->states, e->count, out, per, pp, guard_of(e),
                                                          S(stream), gs));
}

int fill_binomial(erd_engines* e, int32_t* out, int64_t per, int32_t trials, double p, erd::GenState* gs,
                  void* stream) {
    int rc = check_fill(e, out, per, 4);
    if (rc) return rc;
    if (trials < 0 || !(p >= 0.0 && p <= 1.0)) return ERD_ERR_INVALID;
    // BinomialGen constructor (Discrete.h:929-941).
    erd::BinomialParams bp;
    bp.trials = trials;
    bp.p = p;
    bp.small_p = std::min(p, 1 - p);
    bp.mean = trials * bp.small_p;
    bp.g1 = bp.sqrt_v = bp.log_small_p = bp.log_small_q = 0;
    if (trials < 25)
        return hip_status(erd::launch_fill_binomial_small_i32(e->states, e->count, out, per, trials, p,
                                                              S(stream)));
    const int flip = (p == bp.small_p) ? 0 : 1;
    if (bp.mean < 1.0)
        return hip_status(erd::launch_fill_poisson_knuth_i32(e->states, e->count, out, per, bp.mean,
                                                             trials, flip, guard_of(e), S(stream)));
    bp.g1 = std::lgamma(trials + 1);
    bp.sqrt_v = std::sqrt(2 * bp.mean * (1 - bp.small_p));
    bp.log_small_p = std::log(bp.small_p);
    bp.log_small_q = std::log(1 - bp.small_p);
    return hip_status(erd::launch_fill_binomial_reject_i32(e->states, e->count, out, per, bp, guard_of(e),
                                                           S(stream), gs));
}

}  // namespace

extern "C" {

int erd_fill_gamma_f32(erd_engines* e, float* out, int64_t per, float alpha, float beta, void* stream) {
    return fill_gamma(e, out, per, alpha, beta, nullptr, stream);
}

int erd_fill_poisson_i32(erd_engines* e, int32_t* out, int64_t per, double mean, void* stream) {
    return fill_poisson(e, out, per, mean, nullptr, stream);
}

int erd_fill_binomial_i32(erd_engines* e, int32_t* out, int64_t per, int32_t trials, double p,
                          void* stream) {
    return fill_binomial(e, out, per, trials, p, nullptr, stream);
}

// ------------------------------------------------------ generator objects ---
struct erd_generator {
    int32_t dist;
    double a, b;
    int64_t n;
    erd::GenState* state;  // device, one per engine
};

int erd_generator_create(erd_generator** out, int32_t dist, double a, double b, int64_t n_engines,
                         void* stream) {
    if (!out || n_engines <= 0) return ERD_ERR_INVALID;
    *out = nullptr;
    switch (dist) {
        case ERD_DIST_UNIFORM_F32: case ERD_DIST_STD_NORMAL_F32: case ERD_DIST_NORMAL_F32:
        case ERD_DIST_EXPONENTIAL_F32: case ERD_DIST_CAUCHY_F32: case ERD_DIST_STUDENT_T_F32:
        case ERD_DIST_GAMMA_F32: case ERD_DIST_POISSON_I32: case ERD_DIST_BINOMIAL_I32:
        case ERD_DIST_LOGNORMAL_F32: case ERD_DIST_WEIBULL_F32: case ERD_DIST_EXTREME_VALUE_F32: break;
        default: return ERD_ERR_INVALID;
    }
    erd_generator* g = new (std::nothrow) erd_generator{dist, a, b, n_engines, nullptr};
    if (!g) return ERD_ERR_OOM;
    hipError_t err = hipMalloc(&g->state, sizeof(erd::GenState) * (size_t)n_engines);
    if (err == hipSuccess)
        err = hipMemsetAsync(g->state, 0, sizeof(erd::GenState) * (size_t)n_engines, S(stream));
    if (err == hipSuccess) err = hipStreamSynchronize(S(stream));
    if (err != hipSuccess) {
        if (g->state) (void)hipFree(g->state);
        delete g;
        return hip_status(err);
    }
    *out = g;
    return ERD_OK;
}

int erd_generator_destroy(erd_generator* g) {
    if (!g) return ERD_OK;
    hipError_t err = hipDeviceSynchronize();
    if (g->state) (void)hipFree(g->state);
    delete g;
    return hip_status(err);
}

int erd_generator_reset(erd_generator* g, void* stream) {
    if (!g) return ERD_ERR_INVALID;
    return hip_status(hipMemsetAsync(g->state, 0, sizeof(erd::GenState) * (size_t)g->n, S(stream)));
}

int erd_generator_fill(erd_generator* g, erd_engines* e, void* out, int64_t per, void* stream) {
    if (layout_unsupported(e)) return ERD_ERR_UNSUPPORTED;
    if (!g || !e || e->count != g->n) return ERD_ERR_INVALID;
    switch (g->dist) {
        case ERD_DIST_STD_NORMAL_F32:
        case ERD_DIST_NORMAL_F32:
        case ERD_DIST_LOGNORMAL_F32: {
            int rc = check_fill(e, out, per, 8);
            if (rc) return rc;
            const int mode = g->dist == ERD_DIST_NORMAL_F32 ? 1 : g->dist == ERD_DIST_LOGNORMAL_F32 ? 3 : 0;
            return hip_status(erd::launch_fill_normal_gen_f32(e->states, e->count, (float*)out, per, mode,
                                                              (float)g->a, (float)g->b, nullptr, nullptr,
                                                              g->state, S(stream)));
        }
        case ERD_DIST_GAMMA_F32:
            return fill_gamma(e, (float*)out, per, (float)g->a, (float)g->b, g->state, stream);
        case ERD_DIST_POISSON_I32:
            return fill_poisson(e, (int32_t*)out, per, g->a, g->state, stream);
        case ERD_DIST_BINOMIAL_I32: {
            const int32_t trials = (int32_t)g->a;
            if ((double)trials != g->a) return ERD_ERR_INVALID;
            return fill_binomial(e, (int32_t*)out, per, trials, g->b, g->state, stream);
        }
        case ERD_DIST_UNIFORM_F32: return erd_fill_uniform_f32(e, (float*)out, per, (float)g->a, (float)g->b, stream);
        case ERD_DIST_EXPONENTIAL_F32: return erd_fill_exponential_f32(e, (float*)out, per, (float)g->a, stream);
        case ERD_DIST_CAUCHY_F32: return erd_fill_cauchy_f32(e, (float*)out, per, (float)g->a, (float)g->b, stream);
        case ERD_DIST_STUDENT_T_F32: return erd_fill_studentt_f32(e, (float*)out, per, (float)g->a, stream);
        case ERD_DIST_WEIBULL_F32: return erd_fill_weibull_f32(e, (float*)out, per, (float)g->a, (float)g->b, stream);
        case ERD_DIST_EXTREME_VALUE_F32:
            return erd_fill_extreme_value_f32(e, (float*)out, per, (float)g->a, (float)g->b, stream);
        default: return ERD_ERR_INVALID;
    }
}

int erd_fill_binomial_vop_i32(erd_engines* e, int32_t* out, int64_t per, const int32_t* trials,
                              const float* p, void* stream) {
    int rc = check_fill(e, out, per, 4);
    if (rc) return rc;
    if (per > 0 && (!trials || !p)) return ERD_ERR_INVALID;
    return hip_status(erd::launch_fill_binomial_vop_i32(e->states, e->count, out, per,
                                                        erd::Param<int32_t>{trials, 0},
                                                        erd::Param<float>{p, 0.f}, guard_of(e), S(stream)));
}

int erd_fill_binomial_vop_vs_i32(erd_engines* e, int32_t* out, int64_t per, const int32_t* trials, float p,
                                 void* stream) {
    int rc = check_fill(e, out, per, 4);
    if (rc) return rc;
    if (per > 0 && !trials) return ERD_ERR_INVALID;
    return hip_status(erd::launch_fill_binomial_vop_i32(e->states, e->count, out, per,
                                                        erd::Param<int32_t>{trials, 0},
                                                        erd::Param<float>{nullptr, p}, guard_of(e), S(stream)));
}

int erd_fill_binomial_vop_sv_i32(erd_engines* e, int32_t* out, int64_t per, int32_t trials, const float* p,
                                 void* stream) {
    int rc = check_fill(e, out, per, 4);
    if (rc) return rc;
    if (per > 0 && !p) return ERD_ERR_INVALID;
    return hip_status(erd::launch_fill_binomial_vop_i32(e->states, e->count, out, per,
                                                        erd::Param<int32_t>{nullptr, trials},
                                                        erd::Param<float>{p, 0.f}, guard_of(e), S(stream)));
}

int erd_fill_binomial_vop_f32(erd_engines* e, int32_t* out, int64_t per, const int32_t* trials, const float* p,
                              void* stream) {
    int rc = check_fill(e, out, per, 8);
    if (rc) return rc;
    if (per > 0 && (!trials || !p)) return ERD_ERR_INVALID;
    return hip_status(erd::launch_fill_binomial_vop_f32(e->states, e->count, out, per,
                                                        erd::Param<int32_t>{trials, 0},
                                                        erd::Param<float>{p, 0.f}, guard_of(e), S(stream)));
}

int erd_fill_binomial_vop_vs_f32(erd_engines* e, int32_t* out, int64_t per, const int32_t* trials, float p,
                                 void* stream) {
    int rc = check_fill(e, out, per, 8);
    if (rc) return rc;
    if (per > 0 && !trials) return ERD_ERR_INVALID;
    return hip_status(erd::launch_fill_binomial_vop_f32(e->states, e->count, out, per,
                                                        erd::Param<int32_t>{trials, 0},
                                                        erd::Param<float>{nullptr, p}, guard_of(e), S(stream)));
}

int erd_fill_chisquared_f32(erd_engines* e, float* out, int64_t per, float n, void* stream) {
    return erd_fill_gamma_f32(e, out, per, n * 0.5f, 2.0f, stream);
}

int erd_device_alloc(void** ptr, size_t bytes) {
    if (!ptr) return ERD_ERR_INVALID;
    return hip_status(hipMalloc(ptr, bytes));
}
int erd_device_free(void* ptr) { return hip_status(hipFree(ptr)); }
int erd_memcpy_d2h(void* dst, const void* src, size_t bytes, void* stream) {
    hipError_t err = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, S(stream));
    if (err == hipSuccess) err = hipStreamSynchronize(S(stream));
    return hip_status(err);
}
int erd_memcpy_h2d(void* dst, const void* src, size_t bytes, void* stream) {
    hipError_t err = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, S(stream));
    if (err == hipSuccess) err = hipStreamSynchronize(S(stream));
    return hip_status(err);
}
int erd_stream_synchronize(void* stream) { return hip_status(hipStreamSynchronize(S(stream))); }

}  // extern "C"

// ---------------------------------------------------------------- VoP + batched requests ---

extern "C" {

int erd_fill_gamma_vop_f32(erd_engines* e, float* out, int64_t per, const float* alpha,
                           const float* beta, void* stream) {
    int rc = check_fill(e, out, per, 8);
    if (rc) return rc;
    if (per > 0 && (!alpha || !beta)) return ERD_ERR_INVALID;
    return hip_status(erd::launch_fill_gamma_vop_f32(e->states, e->count, out, per, alpha, beta,
                                                     guard_of(e), S(stream)));
}

int erd_fill_poisson_vop_i32(erd_engines* e, int32_t* out, int64_t per, const float* mean,
                             void* stream) {
    int rc = check_fill(e, out, per, 4);
    if (rc) return rc;
    if (per > 0 && !mean) return ERD_ERR_INVALID;
    return hip_status(erd::launch_fill_poisson_vop_i32(e->states, e->count, out, per, mean, guard_of(e),
                                                       S(stream)));
}

// ------------------------------------------------ raw-bit / VoP float fills ---
int erd_fill_bernoulli_f32(erd_engines* e, float* out, int64_t per, double p, void* stream) {
    int rc = check_fill(e, out, per, 8);
    if (rc) return rc;
    if (!(p >= 0 && p <= 1)) return ERD_ERR_INVALID;  // eigen_assert(0 <= _p && _p <= 1)
    const int32_t pi = (int32_t)(uint32_t)(p * 2147483648.0);
    return hip_status(erd::launch_fill_bernoulli_f32(e->states, e->count, out, per, pi, nullptr, S(stream)));
}

int erd_fill_bernoulli_vop_f32(erd_engines* e, float* out, int64_t per, const float* p, void* stream) {
    int rc = check_fill(e, out, per, 8);
    if (rc) return rc;
    if (per > 0 && !p) return ERD_ERR_INVALID;
    return hip_status(erd::launch_fill_bernoulli_f32(e->states, e->count, out, per, 0, p, S(stream)));
}

int erd_fill_balanced_f32(erd_engines* e, float* out, int64_t per, float a, float b, void* stream) {
    int rc = check_fill(e, out, per, 8);
    if (rc) return rc;
    return hip_status(erd::launch_fill_balanced_f32(e->states, e->count, out, per, a, b, nullptr, nullptr,
                                                    S(stream)));
}

int erd_fill_balanced_vop_f32(erd_engines* e, float* out, int64_t per, const float* a, const float* b,
                              void* stream) {
    int rc = check_fill(e, out, per, 8);
    if (rc) return rc;
    if (per > 0 && (!a || !b)) return ERD_ERR_INVALID;
    return hip_status(erd::launch_fill_balanced_f32(e->states, e->count, out, per, 0.f, 0.f, a, b, S(stream)));
}

static int vop_fill(erd_engines* e, int kind, float* out, int64_t per, const float* a, const float* b,
                    bool two, void* stream) {
    int rc = check_fill(e, out, per, 8);
    if (rc) return rc;
    if (per > 0 && (!a || (two && !b))) return ERD_ERR_INVALID;
    return hip_status(erd::launch_fill_vop_f32(e->states, e->count, kind, out, per, a, b, S(stream)));
}

int erd_fill_uniform_vop_f32(erd_engines* e, float* out, int64_t per, const float* a, const float* b,
                             void* stream) {
    return vop_fill(e, erd::VOP_UNIFORM, out, per, a, b, true, stream);
}

int erd_fill_exponential_vop_f32(erd_engines* e, float* out, int64_t per, const float* lambda, void* stream) {
    return vop_fill(e, erd::VOP_EXPONENTIAL, out, per, lambda, nullptr, false, stream);
}

int erd_fill_cauchy_vop_f32(erd_engines* e, float* out, int64_t per, const float* a, const float* b,
                            void* stream) {
    return vop_fill(e, erd::VOP_CAUCHY, out, per, a, b, true, stream);
}

int erd_fill_lognormal_vop_f32(erd_engines* e, float* out, int64_t per, const float* mean, const float* sd,
                               void* stream) {
    return vop_fill(e, erd::VOP_LOGNORMAL, out, per, mean, sd, true, stream);
}

int erd_fill_studentt_vop_f32(erd_engines* e, float* out, int64_t per, const float* n, void* stream) {
    return vop_fill(e, erd::VOP_STUDENT_T, out, per, n, nullptr, false, stream);
}

int erd_fill_weibull_vop_f32(erd_engines* e, float* out, int64_t per, const float* a, const float* b,
                             void* stream) {
    return vop_fill(e, erd::VOP_WEIBULL, out, per, a, b, true, stream);
}

int erd_fill_extreme_value_vop_f32(erd_engines* e, float* out, int64_t per, const float* a, const float* b,
                                   void* stream) {
    return vop_fill(e, erd::VOP_EXTREME_VALUE, out, per, a, b, true, stream);
}

// ------------------------------------------------------ half-draw int32 fills ---
int erd_fill_bernoulli_i32(erd_engines* e, int32_t* out, int64_t per, double p, void* stream) {
    int rc = check_fill(e, out, per, 4);
    if (rc) return rc;
    if (!(p >= 0 && p <= 1)) return ERD_ERR_INVALID;
    const int32_t pi = (int32_t)(uint32_t)(p * 2147483648.0);
    return hip_status(erd::launch_fill_bernoulli_i32(e->states, e->count, out, per, pi, S(stream)));
}

int erd_fill_geometric_i32(erd_engines* e, int32_t* out, int64_t per, double p, void* stream) {
    int rc = check_fill(e, out, per, 4);
    if (rc) return rc;
    if (!(p >= 0 && p <= 1)) return ERD_ERR_INVALID;
    const float rlog_q = (float)(1 / std::log(1 - p));  // GeometricGen's constructor, in double
    return hip_status(erd::launch_fill_geometric_i32(e->states, e->count, out, per, rlog_q, S(stream)));
}

// BetaGen / FisherFGen: one SEG_BETA_F32 segment per engine (the gamma-ratio
// regime alternates two rejection samplers packet by packet).
static int beta_fill(erd_engines* e, float* out, int64_t per, float a, float b, float ratio, void* stream) {
    int rc = check_fill(e, out, per, 8);
    if (rc) return rc;
    if (!(a > 0) || !(b > 0)) return ERD_ERR_INVALID;
    if (per == 0) return ERD_OK;
    std::vector<erd::Segment> segs((size_t)e->count);
    std::vector<int32_t> eng((size_t)e->count);
    for (int64_t g = 0; g < e->count; ++g) {
        erd::Segment& sg = segs[(size_t)g];
        sg = erd::Segment{};
        sg.kind = erd::SEG_BETA_F32;
        sg.count = per;
        sg.out = out + g * per;
        sg.d[0] = a;
        sg.d[1] = b;
        sg.d[2] = ratio;
        eng[(size_t)g] = (int32_t)g;
    }
    return erd::run_segments(e, segs, eng, S(stream));
}

int erd_fill_negative_binomial_i32(erd_engines* e, int32_t* out, int64_t per, int32_t trials, double p,
                                   void* stream) {
    int rc = check_fill(e, out, per, 4);
    if (rc) return rc;
    if (trials <= 0 || !(p > 0 && p <= 1)) return ERD_ERR_INVALID;
    if (per == 0) return ERD_OK;
    std::vector<erd::Segment> segs((size_t)e->count);
    std::vector<int32_t> eng((size_t)e->count);
    for (int64_t g = 0; g < e->count; ++g) {
        erd::Segment& sg = segs[(size_t)g];
        sg = erd::Segment{};
        sg.kind = erd::SEG_NEGATIVE_BINOMIAL_I32;
        sg.count = per;
        sg.out = out + g * per;
        sg.d[0] = (float)trials;             // NegativeBinomialGen: gamma{(float)trials,
        sg.d[1] = (float)((1 - p) / p);      //   (float)((1 - p) / p)}
        eng[(size_t)g] = (int32_t)g;
    }
    return erd::run_segments(e, segs, eng, S(stream));
}

int erd_fill_beta_f32(erd_engines* e, float* out, int64_t per, float a, float b, void* stream) {
    return beta_fill(e, out, per, a, b, 0.0f, stream);
}

int erd_fill_fisher_f_f32(erd_engines* e, float* out, int64_t per, float m, float n, void* stream) {
    const float a = m * 0.5f, b = n * 0.5f;  // FisherFGen: beta{m * 0.5, n * 0.5}
    if (!(a > 0) || !(b > 0)) return ERD_ERR_INVALID;
    return beta_fill(e, out, per, a, b, b / a, stream);
}

int erd_fill_uniform_int_i32(erd_engines* e, int32_t* out, int64_t per, int32_t min, int32_t max, void* stream) {
    int rc = check_fill(e, out, per, 4);
    if (rc) return rc;
    const int64_t pdiff = (int64_t)max - (int64_t)min;
    if (pdiff < 0) return ERD_ERR_INVALID;
    if (pdiff > INT32_MAX) return ERD_ERR_UNSUPPORTED;  // the reference's int32 max - min overflows
    return hip_status(
        erd::launch_fill_uniform_int_i32(e->states, e->count, out, per, min, (uint32_t)pdiff, S(stream)));
}

int erd_fill_randbits_i32(erd_engines* e, int32_t* out, int64_t per, void* stream) {
    int rc = check_fill(e, out, per, 4);
    if (rc) return rc;
    return hip_status(erd::launch_fill_randbits_i32(e->states, e->count, out, per, S(stream)));
}

}  // extern "C"

namespace erd {

// Public request -> internal segment (constructor constants on the host).
int make_segment(const erd_request& r, Segment& s) {
    s = Segment{};
    s.kind = r.dist;
    s.count = r.count;
    s.out = r.out;
    s.in0 = r.in0;
    s.in1 = r.in1;
    if (r.count < 0) return ERD_ERR_INVALID;
    int64_t width = 8;
    switch (r.dist) {
        case ERD_DIST_U64: width = 4; break;
        case ERD_DIST_UNIFORM_F32:
            s.d[0] = (float)r.a; s.d[1] = (float)r.b;
            s.i0 = !((float)r.a == 0.0f && (float)r.b == 1.0f);
            break;
        case ERD_DIST_UNIFORM_F64:
            width = 4;
            s.d[0] = r.a; s.d[1] = r.b;
            s.i0 = !(r.a == 0.0 && r.b == 1.0);
            break;
        case ERD_DIST_STD_NORMAL_F32: break;
        case ERD_DIST_NORMAL_F32: s.d[0] = (float)r.a; s.d[1] = (float)r.b; break;
        case ERD_DIST_NORMAL_VOP_F32:
            if (r.count > 0 && (!r.in0 || !r.in1)) return ERD_ERR_INVALID;
            break;
        case ERD_DIST_EXPONENTIAL_F32: s.d[0] = (float)r.a; break;
        case ERD_DIST_CAUCHY_F32: s.d[0] = (float)r.a; s.d[1] = (float)r.b; break;
        case ERD_DIST_STUDENT_T_F32: s.d[0] = (float)r.a; break;
        case ERD_DIST_GAMMA_F32:
            if (!((float)r.a > 0.0f) || !std::isfinite((float)r.a)) return ERD_ERR_INVALID;
            s.d[0] = (float)r.a; s.d[1] = (float)r.b;
            break;
        case ERD_DIST_POISSON_I32:
            width = 4;
            if (!(r.a >= 0.0) || !std::isfinite(r.a)) return ERD_ERR_INVALID;
            s = seg_poisson((int32_t*)r.out, r.count, r.a);
            break;
        case ERD_DIST_BINOMIAL_I32: {
            width = 4;
            // an integer in [0, INT32_MAX] before the cast (NaN fails the first test)
            if (!(r.a >= 0.0 && r.a <= 2147483647.0) || std::floor(r.a) != r.a || !(r.b >= 0.0 && r.b <= 1.0))
                return ERD_ERR_INVALID;
            const int32_t trials = (int32_t)r.a;
            s = seg_binomial((int32_t*)r.out, r.count, trials, r.b);
            break;
        }
        case ERD_DIST_BINOMIAL_VOP_I32:
        case ERD_DIST_POISSON_VOP_I32:
            width = 4;
            if (r.count > 0 && (!r.in0 || (r.dist == ERD_DIST_BINOMIAL_VOP_I32 && !r.in1)))
                return ERD_ERR_INVALID;
            break;
        case ERD_DIST_GAMMA_VOP_F32:
            if (r.count > 0 && (!r.in0 || !r.in1)) return ERD_ERR_INVALID;
            break;
        case ERD_DIST_DISCARD: width = 1; break;
        case ERD_DIST_LOGNORMAL_F32:
            s.kind = SEG_LOGNORMAL_F32;
            s.d[0] = (float)r.a; s.d[1] = (float)r.b;
            break;
        case ERD_DIST_WEIBULL_F32:
            s.kind = SEG_WEIBULL_F32;
            s.d[0] = 1 / (float)r.a; s.d[1] = (float)r.b;
            break;
        case ERD_DIST_EXTREME_VALUE_F32:
            s.kind = SEG_EXTREME_VALUE_F32;
            s.d[0] = (float)r.a; s.d[1] = (float)r.b;
            break;
        default: return ERD_ERR_INVALID;
    }
    if (r.count % width != 0) return ERD_ERR_INVALID;
    if (r.dist != ERD_DIST_DISCARD && r.count > 0 && !r.out) return ERD_ERR_INVALID;
    return ERD_OK;
}

// Uploads a segment program (segments grouped by engine: consecutive
// segments with the same engine form one group, run in order) and launches
// it. Device buffers come from the stream-ordered allocator and are released
// on the same stream.
int run_segments(erd_engines* e, const std::vector<Segment>& segs, const std::vector<int32_t>& seg_engine,
                 hipStream_t stream, bool p8, int p8_mode) {
    const int64_t n = (int64_t)segs.size();
    if (n == 0) return ERD_OK;
    std::vector<int32_t> engine_of, begin;
    for (int64_t k = 0; k < n; ++k) {
        if (engine_of.empty() || engine_of.back() != seg_engine[(size_t)k]) {
            engine_of.push_back(seg_engine[(size_t)k]);
            begin.push_back((int32_t)k);
        }
    }
    begin.push_back((int32_t)n);
    const int32_t groups = (int32_t)engine_of.size();
    const size_t seg_bytes = sizeof(Segment) * (size_t)n;
    const size_t idx_bytes = sizeof(int32_t) * (size_t)(2 * groups + 1);
    char* dbuf = nullptr;
    hipError_t err = scratch_alloc((void**)&dbuf, seg_bytes + idx_bytes, stream);
    if (err != hipSuccess) return err == hipErrorOutOfMemory ? ERD_ERR_OOM : ERD_ERR_HIP;
    std::vector<char> host(seg_bytes + idx_bytes);
    std::memcpy(host.data(), segs.data(), seg_bytes);
    std::memcpy(host.data() + seg_bytes, engine_of.data(), sizeof(int32_t) * (size_t)groups);
    std::memcpy(host.data() + seg_bytes + sizeof(int32_t) * (size_t)groups, begin.data(),
                sizeof(int32_t) * (size_t)(groups + 1));
    err = upload_async(dbuf, host.data(), host.size(), stream);
    if (err == hipSuccess && p8)
        err = launch_program_p8(e->states, groups, (const int32_t*)(dbuf + seg_bytes),
                                (const int32_t*)(dbuf + seg_bytes) + groups, (const Segment*)dbuf, p8_mode,
                                guard_of(e), stream);
    else if (err == hipSuccess)
        err = launch_program(e->states, groups, (const int32_t*)(dbuf + seg_bytes),
                             (const int32_t*)(dbuf + seg_bytes) + groups, (const Segment*)dbuf,
                             guard_of(e), stream);
    scratch_free(dbuf, seg_bytes + idx_bytes, stream);
    return err == hipSuccess ? ERD_OK : ERD_ERR_HIP;
}

// Groups requests by engine (keeping each engine's order) into a program.
int run_requests(erd_engines* e, const erd_request* reqs, int64_t n, hipStream_t stream) {
    if (!e || !e->states || n < 0 || (n > 0 && !reqs)) return ERD_ERR_INVALID;
    if (n == 0) return ERD_OK;
    std::vector<int64_t> order((size_t)n);
    for (int64_t i = 0; i < n; ++i) {
        if (reqs[i].engine < 0 || reqs[i].engine >= e->count) return ERD_ERR_INVALID;
        order[(size_t)i] = i;
    }
    std::stable_sort(order.begin(), order.end(),
                     [&](int64_t x, int64_t y) { return reqs[x].engine < reqs[y].engine; });
    std::vector<Segment> segs((size_t)n);
    std::vector<int32_t> seg_engine((size_t)n);
    for (int64_t k = 0; k < n; ++k) {
        const erd_request& r = reqs[order[(size_t)k]];
        int rc = make_segment(r, segs[(size_t)k]);
        if (rc) return rc;
        seg_engine[(size_t)k] = r.engine;
    }
    return run_segments(e, segs, seg_engine, stream);
}

}  // namespace erd

extern "C" int erd_run_requests(erd_engines* e, const erd_request* requests, int64_t n, void* stream) {
    if (layout_unsupported(e)) return ERD_ERR_UNSUPPORTED;
    return erd::run_requests(e, requests, n, S(stream));
}
