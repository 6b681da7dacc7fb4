// Host-side launchers for the device kernels (one declaration per kernel
// family). All take device pointers and enqueue on `s`.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace erd {

struct EngineState;

// engine.hip
hipError_t launch_seed(EngineState* st, const uint64_t* d_seeds, int64_t n, hipStream_t s);
hipError_t launch_discard(EngineState* st, int64_t n, int64_t calls, hipStream_t s);

// fill_fixed.hip
hipError_t launch_fill_u64(EngineState* st, int64_t n, uint64_t* out, int64_t per, hipStream_t s);
hipError_t launch_fill_uniform_f32(EngineState* st, int64_t n, float* out, int64_t per, float a,
                                   float b, int affine, hipStream_t s);
hipError_t launch_fill_uniform_f64(EngineState* st, int64_t n, double* out, int64_t per, double a,
                                   double b, int affine, hipStream_t s);
hipError_t launch_fill_normal_f32(EngineState* st, int64_t n, float* out, int64_t per, float mean,
                                  float sd, int affine, hipStream_t s);
hipError_t launch_fill_normal_vop_f32(EngineState* st, int64_t n, float* out, int64_t per,
                                      const float* mean, const float* sd, hipStream_t s);
hipError_t launch_fill_exponential_f32(EngineState* st, int64_t n, float* out, int64_t per,
                                       float lambda, hipStream_t s);
hipError_t launch_fill_cauchy_f32(EngineState* st, int64_t n, float* out, int64_t per, float a,
                                  float b, hipStream_t s);
hipError_t launch_fill_studentt_f32(EngineState* st, int64_t n, float* out, int64_t per, float nu,
                                    hipStream_t s);
hipError_t launch_fill_gamma_fixed_f32(EngineState* st, int64_t n, float* out, int64_t per,
                                       float alpha, float beta, hipStream_t s);

}  // namespace erd
