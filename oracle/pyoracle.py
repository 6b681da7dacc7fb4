"""numpy/ctypes front end of the oracle library (test infrastructure only).

Only tests/, `__graft_entry__.smoke()` and the CPU-baseline leg of bench.py
import this module; the product path never does.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")
STATE_BYTES = 10016

_lib = None

P = C.c_void_p
I64 = C.c_int64


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle not built: {LIB_PATH} (make -C oracle)")
        _lib = C.CDLL(LIB_PATH)
        _lib.orc_state_bytes.restype = C.c_int
        assert _lib.orc_state_bytes() == STATE_BYTES
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


class Engines:
    """Host copy of n Vmt19937_64 engine records (same layout as the device)."""

    def __init__(self, seeds=None, n=None, seed0=0, stride=4, raw=None):
        if raw is not None:
            self.raw = np.array(raw, dtype=np.uint8, copy=True)
            return
        if seeds is None:
            seeds = [seed0 + g * stride for g in range(n)]
        seeds = np.ascontiguousarray(np.asarray(seeds, dtype=np.uint64))
        self.raw = np.zeros(seeds.size * STATE_BYTES, dtype=np.uint8)
        lib().orc_seed(_p(self.raw), _p(seeds), I64(seeds.size))

    def __len__(self):
        return self.raw.size // STATE_BYTES

    def copy(self):
        return Engines(raw=self.raw)

    def subset(self, first, n):
        return Engines(raw=self.raw[first * STATE_BYTES:(first + n) * STATE_BYTES])

    def records(self):
        """Structured view: mt (n,312,4) u64, idx, half_valid, half_cache (n,2)."""
        dt = np.dtype([("mt", "<u8", (312, 4)), ("idx", "<i4"), ("half_valid", "<i4"),
                       ("half_cache", "<u8", (2,)), ("reserved", "<u8")])
        return self.raw.view(dt)


def fill_u64(e, draws, threads=1):
    out = np.empty(len(e) * draws * 4, dtype=np.uint64)
    lib().orc_fill_u64(_p(e.raw), I64(len(e)), _p(out), I64(draws), C.c_int(threads))
    return out


def discard(e, calls, threads=1):
    lib().orc_discard(_p(e.raw), I64(len(e)), I64(calls), C.c_int(threads))


def fill_half_u64(e, halves):
    out = np.empty(len(e) * halves * 2, dtype=np.uint64)
    lib().orc_fill_half_u64(_p(e.raw), I64(len(e)), _p(out), I64(halves))
    return out


def fill_p8_u64(e2, refills):
    assert len(e2) == 2
    out = np.empty(refills * 8, dtype=np.uint64)
    lib().orc_fill_p8_u64(_p(e2.raw), _p(out), I64(refills))
    return out


def _f32(e, per, name, *args, threads=1):
    out = np.empty(len(e) * per, dtype=np.float32)
    getattr(lib(), name)(_p(e.raw), I64(len(e)), _p(out), I64(per), *args, C.c_int(threads))
    return out


def fill_uniform_f32(e, per, a=0.0, b=1.0, threads=1):
    affine = int(not (a == 0.0 and b == 1.0))
    return _f32(e, per, "orc_fill_uniform_f32", C.c_float(a), C.c_float(b), C.c_int(affine),
                threads=threads)


def fill_uniform_f64(e, per, a=0.0, b=1.0, threads=1):
    out = np.empty(len(e) * per, dtype=np.float64)
    affine = int(not (a == 0.0 and b == 1.0))
    lib().orc_fill_uniform_f64(_p(e.raw), I64(len(e)), _p(out), I64(per), C.c_double(a),
                               C.c_double(b), C.c_int(affine), C.c_int(threads))
    return out


def _f64(e, per, name, *args, threads=1):
    out = np.empty(len(e) * per, dtype=np.float64)
    getattr(lib(), name)(_p(e.raw), I64(len(e)), _p(out), I64(per), *args, C.c_int(threads))
    return out


def fill_normal_f64(e, per, mean=None, sd=None, threads=1):
    affine = int(mean is not None or sd is not None)
    return _f64(e, per, "orc_fill_normal_f64", C.c_double(0.0 if mean is None else mean),
                C.c_double(1.0 if sd is None else sd), C.c_int(affine), threads=threads)


def fill_exponential_f64(e, per, lmbda=1.0, threads=1):
    return _f64(e, per, "orc_fill_exponential_f64", C.c_double(lmbda), threads=threads)


def fill_cauchy_f64(e, per, a=0.0, b=1.0, threads=1):
    return _f64(e, per, "orc_fill_cauchy_f64", C.c_double(a), C.c_double(b), threads=threads)


def fill_normal_f32(e, per, mean=None, sd=None, threads=1):
    affine = int(mean is not None or sd is not None)
    return _f32(e, per, "orc_fill_normal_f32", C.c_float(0.0 if mean is None else mean),
                C.c_float(1.0 if sd is None else sd), C.c_int(affine), threads=threads)


def fill_normal_vop_f32(e, mean, sd, threads=1):
    mean = np.ascontiguousarray(mean, dtype=np.float32)
    sd = np.ascontiguousarray(sd, dtype=np.float32)
    per = mean.size // len(e)
    out = np.empty(mean.size, dtype=np.float32)
    lib().orc_fill_normal_vop_f32(_p(e.raw), I64(len(e)), _p(out), I64(per), _p(mean), _p(sd),
                                  C.c_int(threads))
    return out


def fill_exponential_f32(e, per, lmbda=1.0, threads=1):
    return _f32(e, per, "orc_fill_exponential_f32", C.c_float(lmbda), threads=threads)


def fill_lognormal_f32(e, per, mean=0.0, sd=1.0, threads=1):
    return _f32(e, per, "orc_fill_lognormal_f32", C.c_float(mean), C.c_float(sd), threads=threads)


def fill_weibull_f32(e, per, a=1.0, b=1.0, threads=1):
    return _f32(e, per, "orc_fill_weibull_f32", C.c_float(a), C.c_float(b), threads=threads)


def fill_extreme_value_f32(e, per, a=0.0, b=1.0, threads=1):
    return _f32(e, per, "orc_fill_extreme_value_f32", C.c_float(a), C.c_float(b), threads=threads)


def fill_cauchy_f32(e, per, a=0.0, b=1.0, threads=1):
    return _f32(e, per, "orc_fill_cauchy_f32", C.c_float(a), C.c_float(b), threads=threads)


def fill_studentt_f32(e, per, n=1.0, threads=1):
    return _f32(e, per, "orc_fill_studentt_f32", C.c_float(n), threads=threads)


def fill_gamma_f32(e, per, alpha=1.0, beta=1.0, threads=1):
    return _f32(e, per, "orc_fill_gamma_f32", C.c_float(alpha), C.c_float(beta), threads=threads)


def fill_bernoulli_f32(e, per, p=0.5, threads=1):
    return _f32(e, per, "orc_fill_bernoulli_f32", C.c_double(p), None, threads=threads)


def fill_bernoulli_vop_f32(e, p, threads=1):
    p = np.ascontiguousarray(p, dtype=np.float32)
    return _f32(e, p.size // len(e), "orc_fill_bernoulli_f32", C.c_double(0.0), _p(p), threads=threads)


def fill_balanced_f32(e, per, a=-1.0, b=1.0, threads=1):
    return _f32(e, per, "orc_fill_balanced_f32", C.c_float(a), C.c_float(b), None, None, threads=threads)


def fill_balanced_vop_f32(e, a, b, threads=1):
    a = np.ascontiguousarray(a, dtype=np.float32)
    b = np.ascontiguousarray(b, dtype=np.float32)
    return _f32(e, a.size // len(e), "orc_fill_balanced_f32", C.c_float(0), C.c_float(0), _p(a), _p(b),
                threads=threads)


# kind numbers of orc_fill_vop_f32
VOP_KINDS = {"uniform": 0, "exponential": 1, "cauchy": 2, "weibull": 3, "extreme_value": 4,
             "studentt": 5, "lognormal": 6}


def fill_vop_f32(e, kind, a, b=None, threads=1):
    """Vectorize-over-params float fills: a (and b) hold one value per output."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    bb = None if b is None else np.ascontiguousarray(b, dtype=np.float32)
    out = np.empty(a.size, dtype=np.float32)
    lib().orc_fill_vop_f32(_p(e.raw), I64(len(e)), C.c_int(VOP_KINDS[kind]), _p(out), I64(a.size // len(e)),
                           _p(a), None if bb is None else _p(bb), C.c_int(threads))
    return out


def _half_i32(e, per, mode, p, threads):
    out = np.empty(len(e) * per, dtype=np.int32)
    lib().orc_fill_half_i32(_p(e.raw), I64(len(e)), C.c_int(mode), _p(out), I64(per), C.c_double(p),
                            C.c_int(threads))
    return out


def fill_bernoulli_i32(e, per, p=0.5, threads=1):
    return _half_i32(e, per, 0, p, threads)


def fill_geometric_i32(e, per, p=0.5, threads=1):
    return _half_i32(e, per, 1, p, threads)


def fill_randbits_i32(e, per, threads=1):
    return _half_i32(e, per, 2, 0.5, threads)


def fill_uniform_int_i32(e, per, lo, hi, threads=1):
    out = np.empty(len(e) * per, dtype=np.int32)
    lib().orc_fill_uniform_int_i32(_p(e.raw), I64(len(e)), _p(out), I64(per), C.c_int32(lo), C.c_int32(hi),
                                   C.c_int(threads))
    return out


def fill_beta_f32(e, per, a=1.0, b=1.0, threads=1):
    return _f32(e, per, "orc_fill_beta_f32", C.c_float(a), C.c_float(b), C.c_float(0.0), threads=threads)


def fill_fisher_f_f32(e, per, m=1.0, n=1.0, threads=1):
    a, b = np.float32(m) * np.float32(0.5), np.float32(n) * np.float32(0.5)
    return _f32(e, per, "orc_fill_beta_f32", C.c_float(a), C.c_float(b), C.c_float(np.float32(b / a)),
                threads=threads)


def fill_negative_binomial_i32(e, per, trials, p, threads=1):
    out = np.empty(len(e) * per, dtype=np.int32)
    lib().orc_fill_negative_binomial_i32(_p(e.raw), I64(len(e)), _p(out), I64(per), C.c_int32(trials),
                                         C.c_double(p), C.c_int(threads))
    return out


def fill_poisson_i32(e, per, mean, threads=1):
    out = np.empty(len(e) * per, dtype=np.int32)
    lib().orc_fill_poisson_i32(_p(e.raw), I64(len(e)), _p(out), I64(per), C.c_double(mean),
                               C.c_int(threads))
    return out


def fill_binomial_i32(e, per, trials, p, threads=1):
    out = np.empty(len(e) * per, dtype=np.int32)
    lib().orc_fill_binomial_i32(_p(e.raw), I64(len(e)), _p(out), I64(per), C.c_int32(trials),
                                C.c_double(p), C.c_int(threads))
    return out


def fill_binomial_vop_i32(e, trials, p, threads=1):
    trials = np.ascontiguousarray(trials, dtype=np.int32)
    p = np.ascontiguousarray(p, dtype=np.float32)
    per = trials.size // len(e)
    out = np.empty(trials.size, dtype=np.int32)
    lib().orc_fill_binomial_vop_i32(_p(e.raw), I64(len(e)), _p(out), I64(per), _p(trials), _p(p),
                                    C.c_int(threads))
    return out


def fill_binomial_vopf_i32(e, trials, p, threads=1):
    """BinomialVGen<float>: Packet8f lanes on full draws."""
    trials = np.ascontiguousarray(trials, dtype=np.int32)
    p = np.ascontiguousarray(np.broadcast_to(np.float32(p), trials.shape) if np.isscalar(p) else p,
                             dtype=np.float32)
    per = trials.size // len(e)
    out = np.empty(trials.size, dtype=np.int32)
    lib().orc_fill_binomial_vopf_i32(_p(e.raw), I64(len(e)), _p(out), I64(per), _p(trials), _p(p),
                                     C.c_int(threads))
    return out


def _elementwise(name, x, nout=1, dtype=np.float32):
    x = np.ascontiguousarray(x, dtype=dtype)
    outs = [np.empty_like(x) for _ in range(nout)]
    getattr(lib(), name)(_p(x), *[_p(o) for o in outs], I64(x.size))
    return outs[0] if nout == 1 else tuple(outs)


def plog(x):
    return _elementwise("orc_plog", x)


def pexp(x):
    return _elementwise("orc_pexp", x)


def psincos(x):
    return _elementwise("orc_psincos", x, 2)


def plog_f64(x):
    return _elementwise("orc_plog_f64", x, dtype=np.float64)


def psincos_f64(x):
    return _elementwise("orc_psincos_f64", x, 2, dtype=np.float64)


def plgamma(x):
    return _elementwise("orc_plgamma", x)


def bit_to_ur_float(w):
    w = np.ascontiguousarray(w, dtype=np.uint32)
    out = np.empty(w.size, dtype=np.float32)
    lib().orc_bit_to_ur_float(_p(w), _p(out), I64(w.size))
    return out


def compress_stream(W, values, masks, max_packets):
    values = np.ascontiguousarray(values, dtype=np.float32)
    masks = np.ascontiguousarray(masks, dtype=np.uint8)
    n = values.size // W
    out = np.full(max_packets * W, np.nan, dtype=np.float32)
    f = lib().orc_compress_stream
    f.restype = C.c_int64
    used = f(C.c_int(W), _p(values), _p(masks), I64(n), _p(out), I64(max_packets))
    return out.reshape(max_packets, W), used


def fill_gamma_vop_f32(e, alpha, beta, threads=1):
    alpha = np.ascontiguousarray(alpha, dtype=np.float32)
    beta = np.ascontiguousarray(beta, dtype=np.float32)
    per = alpha.size // len(e)
    out = np.empty(alpha.size, dtype=np.float32)
    lib().orc_fill_gamma_vop_f32(_p(e.raw), I64(len(e)), _p(out), I64(per), _p(alpha), _p(beta),
                                 C.c_int(threads))
    return out


def fill_poisson_vop_i32(e, mean, threads=1):
    mean = np.ascontiguousarray(mean, dtype=np.float32)
    per = mean.size // len(e)
    out = np.empty(mean.size, dtype=np.int32)
    lib().orc_fill_poisson_vop_i32(_p(e.raw), I64(len(e)), _p(out), I64(per), _p(mean), C.c_int(threads))
    return out


def alias_table(w):
    w = np.ascontiguousarray(w, dtype=np.float64)
    size = 1 << int(np.ceil(np.log2(w.size)))
    prob = np.zeros(size, dtype=np.int32)
    alias = np.zeros(size, dtype=np.int32)
    bits = lib().orc_alias_table(_p(w), I64(w.size), _p(prob), _p(alias))
    assert bits >= 0
    return prob, alias


def fill_discrete_f_i32(e, per, w):
    """DiscreteGen<int32_t, float>: float CDF for K < 16 (cdf_i = float(c_i /
    c_K), Discrete.h:513-525), else the float alias table."""
    w = np.ascontiguousarray(w, dtype=np.float64)
    out = np.empty(len(e) * per, dtype=np.int32)
    if w.size < 16:
        c = np.cumsum(w)
        cdf = np.ascontiguousarray((c / c[-1]).astype(np.float32))
        lib().orc_fill_discrete_f_i32(_p(e.raw), I64(len(e)), _p(out), I64(per), _p(cdf), C.c_int32(w.size),
                                      None, None, C.c_uint32(0))
        return out
    size = 1 << int(np.ceil(np.log2(w.size)))
    prob = np.zeros(size, dtype=np.float32)
    alias = np.zeros(size, dtype=np.int32)
    bits = lib().orc_alias_table_f32(_p(w), I64(w.size), _p(prob), _p(alias))
    assert bits >= 0
    lib().orc_fill_discrete_f_i32(_p(e.raw), I64(len(e)), _p(out), I64(per), None, C.c_int32(0), _p(prob),
                                  _p(alias), C.c_uint32(size - 1))
    return out


def fill_discrete_i32(e, per, w):
    w = np.ascontiguousarray(w, dtype=np.float64)
    out = np.empty(len(e) * per, dtype=np.int32)
    rc = lib().orc_fill_discrete_i32(_p(e.raw), I64(len(e)), _p(out), I64(per), _p(w), I64(w.size))
    assert rc == 0
    return out


def multinomial(e, trials, weights, samples):
    """Batched MultinomialGen on engines 0..n-1 of e; returns list of (K, samples)."""
    n = len(weights)
    K = np.array([len(w) for w in weights], dtype=np.int64)
    trials = np.broadcast_to(np.asarray(trials, dtype=np.int32), (n,)).copy()
    samples = np.broadcast_to(np.asarray(samples, dtype=np.int64), (n,)).copy()
    w = np.ascontiguousarray(np.concatenate([np.asarray(x, np.float64) for x in weights]))
    out = np.empty(int((K * samples).sum()), dtype=np.int32)
    rc = lib().orc_multinomial(_p(e.raw), I64(n), _p(trials), _p(K), _p(samples), _p(w), _p(out))
    assert rc == 0
    res, off = [], 0
    for k, m in zip(K, samples):
        res.append(out[off:off + k * m].reshape(m, k).T)
        off += k * m
    return res


class Generators:
    """Reused generator objects on n engines (oracle side of erd_generator):
    dist 3 std normal, 4 normal(a, b), 9 gamma(a, b), 10 Poisson(a),
    11 binomial(a, b), 16 lognormal(a, b)."""

    def __init__(self, dist, a, b, n):
        f = lib().orc_gen_create
        f.restype = C.c_void_p
        self.h = C.c_void_p(f(C.c_int(dist), C.c_double(a), C.c_double(b), I64(n)))
        self.dist = dist

    def fill(self, e, per, threads=1):
        dt = np.int32 if self.dist in (10, 11) else np.float32
        out = np.empty(len(e) * per, dtype=dt)
        lib().orc_gen_fill(self.h, _p(e.raw), I64(len(e)), _p(out), I64(per), C.c_int(threads))
        return out

    def __del__(self):
        if getattr(self, "h", None) is not None and _lib is not None:
            _lib.orc_gen_destroy(self.h)
            self.h = None


def std_normal_f32(n, seed=0, threads=1):
    """std::mt19937_64 + std::normal_distribution<float> (the reference
    benchmark's comparator), `threads` contiguous slices."""
    out = np.empty(int(n), dtype=np.float32)
    f = lib().orc_std_normal_f32
    f.restype = None
    f(_p(out), I64(int(n)), C.c_uint64(seed), C.c_int(threads))
    return out


def fill_discrete_d_i32(e, per, w):
    """DiscreteGen<int32_t, double>: double CDF for K < 16 (cdf_i = c_i / c_K,
    Discrete.h:631-645), else AliasMethod<double, int32_t>."""
    w = np.ascontiguousarray(w, dtype=np.float64)
    out = np.empty(len(e) * per, dtype=np.int32)
    f = lib().orc_fill_discrete_d_i32
    f.restype = None
    if w.size < 16:
        c = np.cumsum(w)
        cdf = np.ascontiguousarray(c / c[-1])
        f(_p(e.raw), I64(len(e)), _p(out), I64(per), _p(cdf), C.c_int32(w.size), None, None, C.c_uint32(0))
        return out
    size = 1 << int(np.ceil(np.log2(w.size)))
    prob = np.zeros(size, dtype=np.float64)
    alias = np.zeros(size, dtype=np.int32)
    bits = lib().orc_alias_table_f64(_p(w), I64(w.size), _p(prob), _p(alias))
    assert bits >= 0
    f(_p(e.raw), I64(len(e)), _p(out), I64(per), None, C.c_int32(0), _p(prob), _p(alias), C.c_uint32(size - 1))
    return out


P8_DIST = {"u64": 0, "uniform": 1, "std_normal": 3, "normal": 4, "exponential": 6, "cauchy": 7,
           "student_t": 8, "gamma": 9, "poisson": 10, "binomial": 11, "lognormal": 16, "weibull": 17,
           "extreme_value": 18}


def p8_engines(seeds):
    """Records of P8_mt19937_64 adaptors: adaptor k = Vmt (seed_k), Vmt (seed_k + 4)."""
    return Engines(seeds=[s + d for s in seeds for d in (0, 4)])


def p8_fill(e, dist, per, a=0.0, b=0.0, mode=-1):
    """`dist` with each P8 adaptor of `e` (records 2k, 2k+1) as its engine;
    mode 0 forces operator() draws (the raw stream)."""
    code = P8_DIST[dist]
    dt = np.uint64 if dist == "u64" else np.int32 if dist in ("poisson", "binomial") else np.float32
    out = np.empty((len(e) // 2) * per, dtype=dt)
    f = lib().orc_p8_fill
    f.restype = C.c_int
    rc = f(_p(e.raw), I64(len(e) // 2), C.c_int(code), C.c_int(mode), _p(out), I64(per), C.c_double(a),
           C.c_double(b))
    assert rc == 0
    return out


# Lane layouts (W = 4: SSE2/NEON, W = 16: AVX-512; W = 8 is the AVX2 layout
# of the functions above).
LAYOUT_F32 = {"uniform": 0, "normal": 1, "exponential": 2, "cauchy": 3, "weibull": 4, "extreme_value": 5,
              "studentt": 6}
LAYOUT_F64 = {"uniform": 0, "normal": 1, "exponential": 2, "cauchy": 3}


def fill_u64_layout(e, per, W):
    """per raw u64 per engine at layout W (calls of 2 u64 at W = 4)."""
    out = np.empty(len(e) * per, dtype=np.uint64)
    lib().orc_fill_u64_l(_p(e.raw), I64(len(e)), _p(out), I64(per), C.c_int(W))
    return out


def fill_f32_layout(e, kind, per, W, a=0.0, b=1.0, mode=0):
    out = np.empty(len(e) * per, dtype=np.float32)
    lib().orc_fill_f32_l(_p(e.raw), I64(len(e)), C.c_int(W), C.c_int(LAYOUT_F32[kind]), _p(out), I64(per),
                         C.c_float(a), C.c_float(b), C.c_int(mode))
    return out


def fill_f64_layout(e, kind, per, W, a=0.0, b=1.0, mode=0):
    out = np.empty(len(e) * per, dtype=np.float64)
    lib().orc_fill_f64_l(_p(e.raw), I64(len(e)), C.c_int(W), C.c_int(LAYOUT_F64[kind]), _p(out), I64(per),
                         C.c_double(a), C.c_double(b), C.c_int(mode))
    return out
