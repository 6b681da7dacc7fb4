// Segment programs: a list of fills that run one after another on the same
// engine, in one launch. Used by the multivariate generators (a Wishart or
// Dirichlet draw is a sequence of normal/gamma fills on one engine) and by the
// serving loop (a client's queued requests). Plain data, shared by host and
// device code.
#pragma once
#include <stdint.h>

namespace erd {

enum SegKind : int32_t {
    SEG_U64 = 0,
    SEG_UNIFORM_F32 = 1,     // d[0]=a, d[1]=b, i0=affine
    SEG_UNIFORM_F64 = 2,     // d[0]=a, d[1]=b, i0=affine
    SEG_STD_NORMAL_F32 = 3,
    SEG_NORMAL_F32 = 4,      // d[0]=mean, d[1]=sd
    SEG_NORMAL_VOP_F32 = 5,  // in0=mean, in1=sd
    SEG_EXPONENTIAL_F32 = 6, // d[0]=lambda
    SEG_CAUCHY_F32 = 7,      // d[0]=a, d[1]=b
    SEG_STUDENT_T_F32 = 8,   // d[0]=n
    SEG_GAMMA_F32 = 9,       // d[0]=alpha, d[1]=beta
    SEG_POISSON_I32 = 10,    // d[0..4] = PoissonGen constants
    SEG_BINOMIAL_I32 = 11,   // i0=trials, d[0..7] = BinomialGen constants
    SEG_BINOMIAL_VOP_I32 = 12,  // in0=trials (int32), in1=p (float)
    SEG_GAMMA_VOP_F32 = 13,  // in0=alpha, in1=beta
    SEG_POISSON_VOP_I32 = 14,   // in0=mean (float)
    SEG_DISCARD = 15,        // count = draws to skip
    SEG_DISCRETE_I32 = 16,   // in0=cdf|prob, in1=alias, d[1]=ncdf, d[2]=bitmask
    SEG_MULTINOMIAL_SMALL = 17,  // count=samples, i0=trials, d[0]=K, tables as SEG_DISCRETE_I32
    SEG_MULTINOMIAL_ROWS = 18,   // conditional-binomial rows of one instance: count=samples,
                                 // i0=trials, in0=r (int32 scratch, count), in1=row p (float,
                                 // K-1), d[1]=K; out is the K x count column-major block
    SEG_LOGNORMAL_F32 = 19,      // d[0]=mean, d[1]=sd
    SEG_WEIBULL_F32 = 20,        // d[0]=1/a (float), d[1]=b
    SEG_EXTREME_VALUE_F32 = 21,  // d[0]=a, d[1]=b
    SEG_BETA_F32 = 22,           // d[0]=a, d[1]=b (float), d[2]=FisherF ratio b/a (0: beta)
    SEG_NEGATIVE_BINOMIAL_I32 = 23,  // d[0]=trials (float), d[1]=(1-p)/p (float)
    SEG_KIND_COUNT
};

// How a segment kind takes draws from a P8_mt19937_64 scalar engine (the
// packings are described at P8Feed in erd_ops.hpp): 0 raw draws, 1 Packet8f
// uniform_real packing (three scalars per draw), 2 Packet4i packing (two
// scalars per half draw), -1 not served from a scalar engine.
constexpr int p8_feed_mode(int32_t kind) {
    return kind == SEG_U64 ? 0
         : (kind == SEG_UNIFORM_F32 || kind == SEG_STD_NORMAL_F32 || kind == SEG_NORMAL_F32 ||
            kind == SEG_EXPONENTIAL_F32 || kind == SEG_CAUCHY_F32 || kind == SEG_STUDENT_T_F32 ||
            kind == SEG_GAMMA_F32 || kind == SEG_LOGNORMAL_F32 || kind == SEG_WEIBULL_F32 ||
            kind == SEG_EXTREME_VALUE_F32) ? 1
         : (kind == SEG_POISSON_I32 || kind == SEG_BINOMIAL_I32) ? 2
         : -1;
}

struct Segment {
    int32_t kind;
    int32_t i0;        // integer parameter (trials, affine flag, ...)
    int64_t count;     // outputs (draws for SEG_DISCARD)
    void* out;         // device output (count elements)
    const void* in0;   // per-element parameters (VoP kinds)
    const void* in1;
    double d[8];       // scalar parameters, already in the form the kernel uses
};

}  // namespace erd
