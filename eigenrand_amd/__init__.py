"""eigenrand_amd: EigenRand's sampling hot path on MI355X (gfx950).

Host-side mirror of the reference's generator API over the C ABI
(include/eigenrand_amd.h). Names follow `Eigen::Rand`:

    urng = Vmt19937_64(42)                       # Eigen::Rand::Vmt19937_64 urng{42};
    x = normal(1024, 1024, urng)                 # Rand::normal<MatrixXf>(1024, 1024, urng)
    y = gamma(256, 256, urng, 3.7, 1.0)          # Rand::gamma<MatrixXf>(..., 3.7, 1.0)

A Vmt19937_64 object may hold many engines (a batch of independent
`Vmt19937_64` objects); a fill of `rows x cols` then produces one block per
engine, exactly as if each engine had filled its own matrix. Matrices are
returned column-major like Eigen's default: element (r, c) of block g is at
linear index c*rows + r.

Device memory and streams come from PyTorch (ROCm build); the sampling itself
runs in the native library, which must be built first.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import ErdError, InvalidArgument

__all__ = [
    "Vmt19937_64", "P8_mt19937_64", "Generator", "ErdError", "InvalidArgument", "STATE_BYTES",
    "p8_randbits_u64", "p8_randbits_u32", "p8_uniform_real",
    "randbits_u64", "uniform_real", "uniform_real_f64", "normal", "normal_f64", "normal_vop",
    "exponential_f64", "cauchy_f64",
    "exponential", "cauchy", "student_t", "gamma", "chi_squared", "poisson", "binomial",
    "lognormal", "weibull", "extreme_value",
    "bernoulli", "balanced", "geometric", "randbits_i32", "uniform_int", "beta", "fisher_f", "negative_binomial", "uniform_real_vop", "balanced_vop", "bernoulli_vop",
    "exponential_vop", "cauchy_vop", "lognormal_vop", "student_t_vop", "weibull_vop", "extreme_value_vop",
    "binomial_vop", "gamma_vop", "poisson_vop", "run_requests", "cholesky_batch",
    "mvnormal_batch", "wishart_batch", "invwishart_batch", "dirichlet_batch", "MvNormalGen", "WishartGen",
    "InvWishartGen",
    "MvFactors", "MvNormalBatch", "WishartBatch",
    "DirichletGen", "alias_table", "discrete", "multinomial_batch", "MultinomialGen",
]

STATE_BYTES = 10016  # erd_engine_state_bytes()
STATUS_ITER_CAP = 1  # ERD_STATUS_ITER_CAP
STATUS_BAD_SEGMENT = 2  # ERD_STATUS_BAD_SEGMENT


def _torch():
    import torch
    return torch


def _stream_ptr(stream):
    torch = _torch()
    if stream is None:
        stream = torch.cuda.current_stream()
    return C.c_void_p(stream.cuda_stream)


def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else C.c_void_p(0)


class Vmt19937_64:
    """A set of `n` Vmt19937_64 engines resident in GPU memory.

    Engine g is `Vmt19937_64(seed + g*stride)`: its 4 MT19937-64 lanes are
    seeded seed+g*stride .. seed+g*stride+3 (PacketRandomEngine.h:230-239).
    With the default stride 4, n engines cover lanes seed .. seed+4n-1.

    `layout` is the float packet width of the reference build being matched:
    8 (AVX2, the default), 4 (SSE2/NEON: the 2-lane Vmt19937_64 seeded x0,
    x0+1, i.e. lanes 0-1 of these engines) or 16 (AVX-512: packets of two
    calls). At 4 and 16 only the fills named at erd_engines_set_layout have
    a form; the others raise ErdError(ERD_ERR_UNSUPPORTED).
    """

    default_seed = 5489

    def __init__(self, seed=default_seed, n=1, stride=4, seeds=None, stream=None, layout=8, _handle=None,
                 _parent=None):
        self._parent = _parent
        if _handle is not None:
            self._h = _handle
            return
        h = C.c_void_p()
        if seeds is not None:
            arr = np.ascontiguousarray(np.asarray(seeds, dtype=np.uint64))
            _lib.call("erd_engines_create", C.byref(h), arr.ctypes.data_as(C.c_void_p),
                      int(arr.size), _stream_ptr(stream))
        else:
            _lib.call("erd_engines_create_range", C.byref(h), int(seed) % (1 << 64),
                      int(stride) % (1 << 64), int(n), _stream_ptr(stream))
        self._h = h
        if layout != 8:
            self.set_layout(layout)

    @property
    def layout(self):
        """Float packet width of the fills on this set (4, 8 or 16)."""
        return int(_lib.load().erd_engines_layout(self._h))

    def set_layout(self, layout):
        _lib.call("erd_engines_set_layout", self._h, int(layout))

    @property
    def handle(self):
        return self._h

    def __len__(self):
        return int(_lib.load().erd_engines_count(self._h))

    def __del__(self):
        h = getattr(self, "_h", None)
        lib = getattr(_lib, "_lib", None) if _lib is not None else None  # None at interpreter exit
        if h is not None and lib is not None:
            lib.erd_engines_destroy(h)
            self._h = None

    def view(self, first, n):
        """Engines [first, first+n) of this set, sharing their state."""
        h = C.c_void_p()
        _lib.call("erd_engines_view", C.byref(h), self._h, int(first), int(n))
        return Vmt19937_64(_handle=h, _parent=self)

    def snapshot(self, stream=None):
        """Engine records as a uint8 array (n * STATE_BYTES)."""
        buf = np.empty(len(self) * STATE_BYTES, dtype=np.uint8)
        _lib.call("erd_engines_snapshot", self._h, buf.ctypes.data_as(C.c_void_p), _stream_ptr(stream))
        return buf

    def restore(self, buf, stream=None):
        buf = np.ascontiguousarray(np.asarray(buf, dtype=np.uint8))
        if buf.size != len(self) * STATE_BYTES:
            raise InvalidArgument("erd_engines_restore", 1)
        _lib.call("erd_engines_restore", self._h, buf.ctypes.data_as(C.c_void_p), _stream_ptr(stream))

    def copy(self, stream=None):
        """An independent engine set in the same state (`auto c = urng;`)."""
        out = Vmt19937_64(seed=0, n=len(self), stream=stream, layout=self.layout)
        _lib.call("erd_engines_copy", out._h, self._h, _stream_ptr(stream))
        return out

    def discard(self, calls, stream=None):
        _lib.call("erd_engines_discard", self._h, int(calls), _stream_ptr(stream))

    def status(self, stream=None):
        """Reads and clears the kernel status bits of this set (synchronises
        `stream`): STATUS_ITER_CAP when a sampler gave up at its cap."""
        v = C.c_uint32()
        _lib.call("erd_engines_status", self._h, C.byref(v), _stream_ptr(stream))
        return int(v.value)

    def check(self, stream=None):
        """Raises ErdError(ERD_ERR_ITER_CAP) if a launch on this set gave up at
        an iteration cap since the last status read (the outputs of that
        launch are unspecified); clears the bits."""
        bits = self.status(stream)
        if bits & STATUS_ITER_CAP:
            raise ErdError("sampler iteration cap", 4)
        if bits:
            raise ErdError("malformed request program", 1)

    def set_iteration_caps(self, reject_cap=0, lockstep_cap=0):
        """Caps of later launches on this set: a rejection fill gives up after
        reject_cap candidates without output, a lockstep packet after
        lockstep_cap iterations (0: the defaults 2^22 and 2^20)."""
        _lib.call("erd_engines_set_iteration_caps", self._h, int(reject_cap), int(lockstep_cap))


class P8_mt19937_64(Vmt19937_64):
    """`n` P8_mt19937_64 adaptors (PacketRandomEngine.h:429-517, 600-609).

    Adaptor k = the Vmt19937_64 pair (seed_k, seed_k + 4), so its eight
    64-bit lanes are std::mt19937_64(seed_k .. seed_k + 7) on any ISA. The
    handle holds 2n Vmt engines (len() counts those); `adaptors` is n.
    """

    def __init__(self, seed=Vmt19937_64.default_seed, n=1, stride=8, seeds=None, stream=None):
        if seeds is None:
            seeds = [int(seed) + k * int(stride) for k in range(int(n))]
        arr = np.ascontiguousarray(np.asarray(seeds, dtype=np.uint64))
        h = C.c_void_p()
        _lib.call("erd_engines_create_p8", C.byref(h), arr.ctypes.data_as(C.c_void_p), int(arr.size),
                  _stream_ptr(stream))
        super().__init__(_handle=h)

    @property
    def adaptors(self):
        return len(self) // 2


def _p8_fill(name, count, urng, dtype, out, stream):
    torch = _torch()
    n = len(urng) // 2
    if out is None:
        out = torch.empty(n * int(count), dtype=dtype, device="cuda")
    elif out.dtype != dtype or not out.is_cuda or not out.is_contiguous() or out.numel() != n * count:
        raise InvalidArgument("output tensor", 1)
    _lib.call(name, urng.handle, _ptr(out), int(count), _stream_ptr(stream))
    return out.view(n, int(count))


def p8_randbits_u64(count, urng, out=None, stream=None):
    """`count` calls of P8_mt19937_64::operator() per adaptor; (n, count) uint64."""
    return _p8_fill("erd_p8_fill_u64", count, urng, _torch().uint64, out, stream)


def p8_randbits_u32(count, urng, out=None, stream=None):
    """`count` calls of P8_mt19937_64_32::operator() per adaptor; (n, count) uint32."""
    return _p8_fill("erd_p8_fill_u32", count, urng, _torch().uint32, out, stream)


def p8_uniform_real(count, urng, out=None, stream=None):
    """`count` calls of the adaptor's uniform_real() per adaptor; (n, count) float32."""
    return _p8_fill("erd_p8_fill_uniform_f32", count, urng, _torch().float32, out, stream)


P8_DISTS = ("uniform", "std_normal", "normal", "exponential", "cauchy", "student_t", "gamma",
            "lognormal", "weibull", "extreme_value", "poisson", "binomial")


def p8_fill(dist, count, urng, a=0.0, b=0.0, out=None, stream=None):
    """`Rand::dist<ArrayXf or ArrayXi>(count, 1, p8_k, a, b)` with each
    P8_mt19937_64 adaptor k of `urng` as the engine (test/test.cpp:117 and
    on): the packet code packs the adaptor's scalar draws into 256-bit words
    (RawbitsMaker, arch/AVX/RandUtils.h:59-144) in Clang's argument order.
    (n, count) float32, or int32 for poisson/binomial. a, b as in
    `Request` (binomial: a = trials, b = p)."""
    torch = _torch()
    if dist not in P8_DISTS:
        raise ValueError(f"p8_fill: {dist!r} is not drawn from a P8 adaptor; one of {P8_DISTS}")
    dt = torch.int32 if dist in ("poisson", "binomial") else torch.float32
    n = len(urng) // 2
    if out is None:
        out = torch.empty(n * int(count), dtype=dt, device="cuda")
    elif out.dtype != dt or not out.is_cuda or not out.is_contiguous() or out.numel() != n * count:
        raise InvalidArgument("output tensor", 1)
    _lib.call("erd_p8_fill_dist", urng.handle, _lib.DIST[dist], _ptr(out), int(count), float(a), float(b),
              _stream_ptr(stream))
    return out.view(n, int(count))


class Generator:
    """A reused generator object on `n_engines` engines: the GPU form of
    `auto gen = Rand::makeXxxGen(a, b)` followed by several
    gen.generate(per_engine, 1, engine_g) calls. It keeps the normal
    generator's cached cos packet and the rejection samplers' CompressMask
    leftovers between calls (Basic.h:268-342), so consecutive fills equal one
    larger fill. dist: "uniform", "std_normal", "normal", "exponential",
    "cauchy", "student_t", "gamma", "poisson" or "binomial"."""

    def __init__(self, dist, a=0.0, b=0.0, n_engines=1, stream=None):
        self.dist = dist
        kind = _lib.DIST[dist]
        h = C.c_void_p()
        _lib.call("erd_generator_create", C.byref(h), kind, float(a), float(b), int(n_engines),
                  _stream_ptr(stream))
        self._h = h
        self.n = int(n_engines)

    def generate(self, per_engine, urng, out=None, stream=None):
        """Flat output: engine g's block is out[g*per_engine:(g+1)*per_engine]."""
        torch = _torch()
        dtype = torch.int32 if self.dist in ("poisson", "binomial") else torch.float32
        buf = _alloc(urng, int(per_engine), dtype, out)
        _lib.call("erd_generator_fill", self._h, urng.handle, _ptr(buf), int(per_engine), _stream_ptr(stream))
        return buf

    def reset(self, stream=None):
        _lib.call("erd_generator_reset", self._h, _stream_ptr(stream))

    def __del__(self):
        h = getattr(self, "_h", None)
        lib = getattr(_lib, "_lib", None) if _lib is not None else None  # None at interpreter exit
        if h is not None and lib is not None:
            lib.erd_generator_destroy(h)
            self._h = None


def _alloc(urng, per_engine, dtype, out):
    torch = _torch()
    n = len(urng)
    if out is None:
        out = torch.empty(n * per_engine, dtype=dtype, device="cuda")
    else:
        if out.dtype != dtype or not out.is_cuda or not out.is_contiguous() or out.numel() != n * per_engine:
            raise InvalidArgument("output tensor", 1)
    return out


def _param(t, dtype, name):
    """A per-element parameter tensor, checked before any launch: it must be
    a torch tensor of exactly `dtype` on the current CUDA device (where the
    output lives). A host tensor would hand the kernel a host pointer, a
    narrower dtype would make it read past the allocation, and a wider one
    would be misread."""
    torch = _torch()
    if not isinstance(t, torch.Tensor) or t.dtype != dtype or not t.is_cuda \
            or t.device.index != torch.cuda.current_device():
        raise InvalidArgument(name, 1)
    return t.contiguous()


def _shape(flat, urng, rows, cols):
    n = len(urng)
    if n == 1:
        return flat.view(cols, rows).t()
    return flat.view(n, cols, rows).transpose(1, 2)


def _fill(name, urng, rows, cols, dtype, out, stream, *params):
    per = int(rows) * int(cols)
    buf = _alloc(urng, per, dtype, out)
    _lib.call(name, urng.handle, _ptr(buf), per, *params, _stream_ptr(stream))
    return _shape(buf, urng, rows, cols) if out is None else buf


def randbits_u64(draws, urng, out=None, stream=None):
    """`draws` calls of Vmt19937_64::operator(); returns (n, 4*draws) uint64
    ((n, 2*draws) at layout 4, whose calls are two lanes wide)."""
    torch = _torch()
    per = (2 if urng.layout == 4 else 4) * int(draws)
    buf = _alloc(urng, per, torch.uint64, out)
    _lib.call("erd_fill_u64", urng.handle, _ptr(buf), per, _stream_ptr(stream))
    return buf.view(len(urng), per)


def uniform_real(rows, cols, urng, min=0.0, max=1.0, out=None, stream=None):
    """Rand::uniformReal<MatrixXf>(rows, cols, urng[, min, max])."""
    return _fill("erd_fill_uniform_f32", urng, rows, cols, _torch().float32, out, stream,
                 float(min), float(max))


def uniform_real_f64(rows, cols, urng, min=0.0, max=1.0, out=None, stream=None):
    """Rand::uniformReal<MatrixXd>(rows, cols, urng[, min, max])."""
    return _fill("erd_fill_uniform_f64", urng, rows, cols, _torch().float64, out, stream,
                 float(min), float(max))


def normal(rows, cols, urng, mean=None, stdev=None, out=None, stream=None, precision="exact"):
    """Rand::normal<MatrixXf>(rows, cols, urng) (StdNormalGen) or, with
    mean/stdev, Rand::normal<MatrixXf>(rows, cols, urng, mean, stdev) (NormalGen).

    precision="exact": the reference's float op sequence, bit for bit.
    precision="fast": the same engine draws, pairing and output order with the
    hardware log2/sqrt/sin/cos; each value within 4 ulp + 2^-20 * r of the
    exact mode's (r = sqrt(-2 ln u1) <= 5.65)."""
    torch = _torch()
    if precision not in ("exact", "fast"):
        raise InvalidArgument("normal: precision", 1)
    sfx = "_fast" if precision == "fast" else ""
    if mean is None and stdev is None:
        return _fill("erd_fill_std_normal_f32" + sfx, urng, rows, cols, torch.float32, out, stream)
    return _fill("erd_fill_normal_f32" + sfx, urng, rows, cols, torch.float32, out, stream,
                 float(0.0 if mean is None else mean), float(1.0 if stdev is None else stdev))


def store_ceiling(n_blocks, per_block, out, stream=None):
    """Write-only 16-byte stores of 1.0f into n_blocks contiguous blocks of
    per_block floats (one 64-thread workgroup per block): the HBM-write
    ceiling the fills are measured against. No reference counterpart."""
    torch = _torch()
    if out.dtype != torch.float32 or not out.is_cuda or not out.is_contiguous() \
            or out.numel() < int(n_blocks) * int(per_block):
        raise InvalidArgument("store_ceiling: output tensor", 1)
    _lib.call("erd_store_ceiling_f32", _ptr(out), int(n_blocks), int(per_block), _stream_ptr(stream))
    return out


def normal_f64(rows, cols, urng, mean=None, stdev=None, out=None, stream=None):
    """Rand::normal<MatrixXd>(rows, cols, urng[, mean, stdev]): StdNormalGen /
    NormalGen at Packet4d (four doubles per engine draw)."""
    torch = _torch()
    if mean is None and stdev is None:
        return _fill("erd_fill_std_normal_f64", urng, rows, cols, torch.float64, out, stream)
    return _fill("erd_fill_normal_f64", urng, rows, cols, torch.float64, out, stream,
                 float(0.0 if mean is None else mean), float(1.0 if stdev is None else stdev))


def normal_vop(urng, mean, stdev, out=None, stream=None):
    """Rand::normal(urng, mean_array, stdev_array) (NormalVGen); mean/stdev are
    CUDA float32 tensors of n*per_engine elements (engine g uses its block).
    One of them may be a Python number: the reference's mixed overloads
    normal(urng, a, b_scalar) / normal(urng, a_scalar, b) (NormalExp.h:
    1304-1335)."""
    torch = _torch()
    if isinstance(stdev, (int, float)) and not isinstance(mean, (int, float)):
        mean = _param(mean, torch.float32, "erd_fill_normal_vop_vs_f32: mean")
        per = _vop_per(urng, mean.numel(), "erd_fill_normal_vop_vs_f32")
        buf = _alloc(urng, per, torch.float32, out)
        _lib.call("erd_fill_normal_vop_vs_f32", urng.handle, _ptr(buf), per, _ptr(mean), float(stdev),
                  _stream_ptr(stream))
        return buf
    if isinstance(mean, (int, float)) and not isinstance(stdev, (int, float)):
        stdev = _param(stdev, torch.float32, "erd_fill_normal_vop_sv_f32: stdev")
        per = _vop_per(urng, stdev.numel(), "erd_fill_normal_vop_sv_f32")
        buf = _alloc(urng, per, torch.float32, out)
        _lib.call("erd_fill_normal_vop_sv_f32", urng.handle, _ptr(buf), per, float(mean), _ptr(stdev),
                  _stream_ptr(stream))
        return buf
    mean = _param(mean, torch.float32, "erd_fill_normal_vop_f32: mean")
    stdev = _param(stdev, torch.float32, "erd_fill_normal_vop_f32: stdev")
    if mean.numel() != stdev.numel() or mean.numel() % len(urng):
        raise InvalidArgument("erd_fill_normal_vop_f32", 1)
    per = mean.numel() // len(urng)
    buf = _alloc(urng, per, torch.float32, out)
    _lib.call("erd_fill_normal_vop_f32", urng.handle, _ptr(buf), per, _ptr(mean), _ptr(stdev),
              _stream_ptr(stream))
    return buf


def _vop_per(urng, numel, name):
    if numel % len(urng):
        raise InvalidArgument(name, 1)
    return numel // len(urng)


def exponential(rows, cols, urng, lmbda=1.0, out=None, stream=None):
    """Rand::exponential<MatrixXf>(rows, cols, urng, lambda)."""
    return _fill("erd_fill_exponential_f32", urng, rows, cols, _torch().float32, out, stream,
                 float(lmbda))


def exponential_f64(rows, cols, urng, lmbda=1.0, out=None, stream=None):
    """Rand::exponential<MatrixXd>(rows, cols, urng, lambda)."""
    return _fill("erd_fill_exponential_f64", urng, rows, cols, _torch().float64, out, stream,
                 float(lmbda))


def lognormal(rows, cols, urng, mean=0.0, stdev=1.0, out=None, stream=None):
    """Rand::lognormal<MatrixXf>(rows, cols, urng, mean, stdev) (LognormalGen)."""
    return _fill("erd_fill_lognormal_f32", urng, rows, cols, _torch().float32, out, stream,
                 float(mean), float(stdev))


def weibull(rows, cols, urng, a=1.0, b=1.0, out=None, stream=None):
    """Rand::weibull<MatrixXf>(rows, cols, urng, a, b) (WeibullGen)."""
    return _fill("erd_fill_weibull_f32", urng, rows, cols, _torch().float32, out, stream, float(a), float(b))


def extreme_value(rows, cols, urng, a=0.0, b=1.0, out=None, stream=None):
    """Rand::extremeValue<MatrixXf>(rows, cols, urng, a, b) (ExtremeValueGen)."""
    return _fill("erd_fill_extreme_value_f32", urng, rows, cols, _torch().float32, out, stream,
                 float(a), float(b))


def cauchy(rows, cols, urng, a=0.0, b=1.0, out=None, stream=None):
    """Rand::cauchy<MatrixXf>(rows, cols, urng, a, b)."""
    return _fill("erd_fill_cauchy_f32", urng, rows, cols, _torch().float32, out, stream,
                 float(a), float(b))


def cauchy_f64(rows, cols, urng, a=0.0, b=1.0, out=None, stream=None):
    """Rand::cauchy<MatrixXd>(rows, cols, urng, a, b)."""
    return _fill("erd_fill_cauchy_f64", urng, rows, cols, _torch().float64, out, stream,
                 float(a), float(b))


def student_t(rows, cols, urng, n=1.0, out=None, stream=None):
    """Rand::studentT<MatrixXf>(rows, cols, urng, n)."""
    return _fill("erd_fill_studentt_f32", urng, rows, cols, _torch().float32, out, stream, float(n))


def gamma(rows, cols, urng, alpha=1.0, beta=1.0, out=None, stream=None):
    """Rand::gamma<MatrixXf>(rows, cols, urng, alpha, beta)."""
    return _fill("erd_fill_gamma_f32", urng, rows, cols, _torch().float32, out, stream,
                 float(alpha), float(beta))


def chi_squared(rows, cols, urng, n=1.0, out=None, stream=None):
    """Rand::chiSquared<MatrixXf>(rows, cols, urng, n)."""
    return _fill("erd_fill_chisquared_f32", urng, rows, cols, _torch().float32, out, stream, float(n))


def poisson(rows, cols, urng, mean=1.0, out=None, stream=None):
    """Rand::poisson<MatrixXi>(rows, cols, urng, mean) (int32)."""
    return _fill("erd_fill_poisson_i32", urng, rows, cols, _torch().int32, out, stream, float(mean))


def binomial(rows, cols, urng, trials=1, p=0.5, out=None, stream=None):
    """Rand::binomial<MatrixXi>(rows, cols, urng, trials, p) (int32)."""
    return _fill("erd_fill_binomial_i32", urng, rows, cols, _torch().int32, out, stream, int(trials),
                 float(p))


def bernoulli(rows, cols, urng, p=0.5, out=None, stream=None, dtype=None):
    """Rand::bernoulli<MatrixXf>(rows, cols, urng, p) (BernoulliGen): float32
    1/0 from full draws; dtype=torch.int32 gives the int32 (Packet4i, half
    draw) form."""
    torch = _torch()
    if dtype is torch.int32:
        return _fill("erd_fill_bernoulli_i32", urng, rows, cols, torch.int32, out, stream, float(p))
    return _fill("erd_fill_bernoulli_f32", urng, rows, cols, torch.float32, out, stream, float(p))


def balanced(rows, cols, urng, a=-1.0, b=1.0, out=None, stream=None):
    """Rand::balanced<MatrixXf>(rows, cols, urng[, a, b]) (BalancedGen /
    Balanced2Gen): reals in [a, b]."""
    return _fill("erd_fill_balanced_f32", urng, rows, cols, _torch().float32, out, stream, float(a), float(b))


def beta(rows, cols, urng, a=1.0, b=1.0, out=None, stream=None):
    """Rand::beta<MatrixXf>(rows, cols, urng, a, b) (BetaGen)."""
    return _fill("erd_fill_beta_f32", urng, rows, cols, _torch().float32, out, stream, float(a), float(b))


def fisher_f(rows, cols, urng, m=1.0, n=1.0, out=None, stream=None):
    """Rand::fisherF<MatrixXf>(rows, cols, urng, m, n) (FisherFGen)."""
    return _fill("erd_fill_fisher_f_f32", urng, rows, cols, _torch().float32, out, stream, float(m), float(n))


def negative_binomial(rows, cols, urng, trials=1, p=0.5, out=None, stream=None):
    """Rand::negativeBinomial<MatrixXi>(rows, cols, urng, trials, p) (int32)."""
    return _fill("erd_fill_negative_binomial_i32", urng, rows, cols, _torch().int32, out, stream, int(trials),
                 float(p))


def geometric(rows, cols, urng, p=0.5, out=None, stream=None):
    """Rand::geometric<MatrixXi>(rows, cols, urng, p) (GeometricGen, int32)."""
    return _fill("erd_fill_geometric_i32", urng, rows, cols, _torch().int32, out, stream, float(p))


def uniform_int(rows, cols, urng, min=0, max=0, out=None, stream=None):
    """Rand::uniformInt<MatrixXi>(rows, cols, urng, min, max) (UniformIntGen, int32)."""
    return _fill("erd_fill_uniform_int_i32", urng, rows, cols, _torch().int32, out, stream, int(min), int(max))


def randbits_i32(rows, cols, urng, out=None, stream=None):
    """Rand::randBits<MatrixXi>(rows, cols, urng) (RandbitsGen<int32_t>)."""
    return _fill("erd_fill_randbits_i32", urng, rows, cols, _torch().int32, out, stream)


def _vop(name, urng, params, out, stream):
    """Vectorize-over-params fill: params are float32 CUDA tensors of
    n*per_engine elements (engine g uses its block)."""
    torch = _torch()
    params = [_param(p, torch.float32, name) for p in params]
    if any(p.numel() != params[0].numel() for p in params) or params[0].numel() % len(urng):
        raise InvalidArgument(name, 1)
    per = params[0].numel() // len(urng)
    buf = _alloc(urng, per, torch.float32, out)
    _lib.call(name, urng.handle, _ptr(buf), per, *[_ptr(p) for p in params], _stream_ptr(stream))
    return buf


def uniform_real_vop(urng, a, b, out=None, stream=None):
    """Rand::uniformReal(urng, a, b) with per-element bounds (UniformRealVGen)."""
    return _vop("erd_fill_uniform_vop_f32", urng, [a, b], out, stream)


def balanced_vop(urng, a, b, out=None, stream=None):
    """Rand::balanced(urng, a, b) with per-element bounds (BalancedVGen)."""
    return _vop("erd_fill_balanced_vop_f32", urng, [a, b], out, stream)


def bernoulli_vop(urng, p, out=None, stream=None):
    """Rand::bernoulli(urng, p_array) (BernoulliVGen<float>)."""
    return _vop("erd_fill_bernoulli_vop_f32", urng, [p], out, stream)


def exponential_vop(urng, lmbda, out=None, stream=None):
    """Rand::exponential(urng, lambda_array) (ExponentialVGen)."""
    return _vop("erd_fill_exponential_vop_f32", urng, [lmbda], out, stream)


def cauchy_vop(urng, a, b, out=None, stream=None):
    """Rand::cauchy(urng, a, b) with per-element parameters (CauchyVGen)."""
    return _vop("erd_fill_cauchy_vop_f32", urng, [a, b], out, stream)


def lognormal_vop(urng, mean, stdev, out=None, stream=None):
    """Rand::lognormal(urng, mean, stdev) with per-element parameters (LognormalVGen)."""
    return _vop("erd_fill_lognormal_vop_f32", urng, [mean, stdev], out, stream)


def student_t_vop(urng, n, out=None, stream=None):
    """Rand::studentT(urng, n_array) (StudentTVGen)."""
    return _vop("erd_fill_studentt_vop_f32", urng, [n], out, stream)


def weibull_vop(urng, a, b, out=None, stream=None):
    """Rand::weibull(urng, a, b) with per-element parameters (WeibullVGen)."""
    return _vop("erd_fill_weibull_vop_f32", urng, [a, b], out, stream)


def extreme_value_vop(urng, a, b, out=None, stream=None):
    """Rand::extremeValue(urng, a, b) with per-element parameters (ExtremeValueVGen)."""
    return _vop("erd_fill_extreme_value_vop_f32", urng, [a, b], out, stream)


def binomial_vop(urng, trials, p, out=None, stream=None):
    """Rand::binomial(urng, trials_array, p_array) (BinomialVGen<int32_t>);
    trials int32 and p float32 CUDA tensors of n*per_engine elements. One of
    them may be a Python number: the mixed overloads binomial(urng, trials, p)
    with a scalar p or scalar trials (Discrete.h:1852-1897)."""
    torch = _torch()
    if isinstance(p, (int, float)) and not isinstance(trials, (int, float)):
        trials = _param(trials, torch.int32, "erd_fill_binomial_vop_vs_i32: trials")
        per = _vop_per(urng, trials.numel(), "erd_fill_binomial_vop_vs_i32")
        buf = _alloc(urng, per, torch.int32, out)
        _lib.call("erd_fill_binomial_vop_vs_i32", urng.handle, _ptr(buf), per, _ptr(trials), float(p),
                  _stream_ptr(stream))
        return buf
    if isinstance(trials, (int, np.integer)) and not isinstance(p, (int, float)):
        p = _param(p, torch.float32, "erd_fill_binomial_vop_sv_i32: p")
        per = _vop_per(urng, p.numel(), "erd_fill_binomial_vop_sv_i32")
        buf = _alloc(urng, per, torch.int32, out)
        _lib.call("erd_fill_binomial_vop_sv_i32", urng.handle, _ptr(buf), per, int(trials), _ptr(p),
                  _stream_ptr(stream))
        return buf
    trials = _param(trials, torch.int32, "erd_fill_binomial_vop_i32: trials")
    p = _param(p, torch.float32, "erd_fill_binomial_vop_i32: p")
    if trials.numel() != p.numel() or trials.numel() % len(urng):
        raise InvalidArgument("erd_fill_binomial_vop_i32", 1)
    per = trials.numel() // len(urng)
    buf = _alloc(urng, per, torch.int32, out)
    _lib.call("erd_fill_binomial_vop_i32", urng.handle, _ptr(buf), per, _ptr(trials), _ptr(p),
              _stream_ptr(stream))
    return buf


def binomial_vop_float(urng, trials, p, out=None, stream=None):
    """Rand::binomial(urng, trials, p) with float-typed trials
    (BinomialVGen<float>, Discrete.h:1196-1331): Packet8f lanes on full draws.
    The reference keeps trials and results as int32 bit patterns in float
    lanes; here trials is an int32 tensor and the result is int32. p may be a
    Python number (the scalar-p form MultinomialGen uses)."""
    torch = _torch()
    trials = _param(trials, torch.int32, "erd_fill_binomial_vop_f32: trials")
    per = _vop_per(urng, trials.numel(), "erd_fill_binomial_vop_f32")
    buf = _alloc(urng, per, torch.int32, out)
    if isinstance(p, (int, float)):
        _lib.call("erd_fill_binomial_vop_vs_f32", urng.handle, _ptr(buf), per, _ptr(trials), float(p),
                  _stream_ptr(stream))
        return buf
    p = _param(p, torch.float32, "erd_fill_binomial_vop_f32: p")
    if p.numel() != trials.numel():
        raise InvalidArgument("erd_fill_binomial_vop_f32", 1)
    _lib.call("erd_fill_binomial_vop_f32", urng.handle, _ptr(buf), per, _ptr(trials), _ptr(p),
              _stream_ptr(stream))
    return buf


def gamma_vop(urng, alpha, beta, out=None, stream=None):
    """Per-element gamma(alpha[i], beta[i]) (float32 CUDA tensors); the
    lockstep vectorize-over-params design of BinomialVGen applied to GammaGen."""
    torch = _torch()
    alpha = _param(alpha, torch.float32, "erd_fill_gamma_vop_f32: alpha")
    beta = _param(beta, torch.float32, "erd_fill_gamma_vop_f32: beta")
    if alpha.numel() != beta.numel() or alpha.numel() % len(urng):
        raise InvalidArgument("erd_fill_gamma_vop_f32", 1)
    per = alpha.numel() // len(urng)
    buf = _alloc(urng, per, torch.float32, out)
    _lib.call("erd_fill_gamma_vop_f32", urng.handle, _ptr(buf), per, _ptr(alpha), _ptr(beta),
              _stream_ptr(stream))
    return buf


def poisson_vop(urng, mean, out=None, stream=None):
    """Per-element poisson(mean[i]) (float32 CUDA tensor) -> int32."""
    torch = _torch()
    mean = _param(mean, torch.float32, "erd_fill_poisson_vop_i32: mean")
    if mean.numel() % len(urng):
        raise InvalidArgument("erd_fill_poisson_vop_i32", 1)
    per = mean.numel() // len(urng)
    buf = _alloc(urng, per, torch.int32, out)
    _lib.call("erd_fill_poisson_vop_i32", urng.handle, _ptr(buf), per, _ptr(mean), _stream_ptr(stream))
    return buf


def run_requests(urng, requests, stream=None):
    """Run a batch of heterogeneous fills (erd_run_requests). `requests` is a
    list of dicts: dist (name), engine, count, out (tensor), optional in0/in1
    (tensors), a, b."""
    torch = _torch()
    arr = (_lib.Request * len(requests))()
    for i, r in enumerate(requests):
        for key in ("out", "in0", "in1"):
            t = r.get(key)
            if t is not None and (not isinstance(t, torch.Tensor) or not t.is_cuda
                                  or t.device.index != torch.cuda.current_device()
                                  or not t.is_contiguous()):
                raise InvalidArgument(f"erd_run_requests: request {i} {key}", 1)
        arr[i].dist = _lib.DIST[r["dist"]]
        arr[i].engine = int(r.get("engine", 0))
        arr[i].count = int(r["count"])
        arr[i].out = r["out"].data_ptr() if r.get("out") is not None else None
        arr[i].in0 = r["in0"].data_ptr() if r.get("in0") is not None else None
        arr[i].in1 = r["in1"].data_ptr() if r.get("in1") is not None else None
        arr[i].a = float(r.get("a", 0.0))
        arr[i].b = float(r.get("b", 0.0))
    _lib.call("erd_run_requests", urng.handle, C.cast(arr, C.c_void_p), len(requests), _stream_ptr(stream))


# ------------------------------------------------------------ multivariate ---
def _concat_f64(mats):
    import torch
    flat = [np.asarray(m, dtype=np.float64).reshape(-1, order="F") for m in mats]
    return torch.from_numpy(np.concatenate(flat)).cuda()


def cholesky_batch(mats, stream=None):
    """FP64 Cholesky of each matrix on the GPU; returns (list of L, info)."""
    torch = _torch()
    dims = np.array([np.asarray(m).shape[0] for m in mats], dtype=np.int32)
    a = _concat_f64(mats)
    info = np.zeros(len(mats), dtype=np.int32)
    rc = _lib.load().erd_cholesky_f64_batch(len(mats), dims.ctypes.data_as(C.c_void_p), _ptr(a),
                                            info.ctypes.data_as(C.c_void_p), _stream_ptr(stream))
    torch.cuda.synchronize()
    out, off = [], 0
    flat = a.cpu().numpy()
    for d in dims:
        out.append(flat[off:off + d * d].reshape(d, d, order="F"))
        off += d * d
    return out, info, rc


def mvnormal_batch(urng, covs, samples, means=None, return_z=False, stream=None):
    """Batched MvNormalGen<float>(mean_i, cov_i).generate(engine_i, samples_i).

    Instance i uses engine i of `urng`. Returns a list of (d_i, samples_i)
    float32 CUDA tensors (and the standard normals Z_i if return_z)."""
    torch = _torch()
    n = len(covs)
    dims = np.array([np.asarray(c).shape[0] for c in covs], dtype=np.int32)
    samples = np.broadcast_to(np.asarray(samples, dtype=np.int64), (n,)).copy()
    cov = _concat_f64(covs)
    mean = None
    if means is not None:
        mean = torch.from_numpy(np.concatenate([np.asarray(m, np.float32).ravel() for m in means])).cuda()
    total = int((dims.astype(np.int64) * samples).sum())
    out = torch.empty(total, dtype=torch.float32, device="cuda")
    z = torch.empty(total, dtype=torch.float32, device="cuda") if return_z else None
    _lib.call("erd_mvnormal_batch", urng.handle, n, dims.ctypes.data_as(C.c_void_p),
              samples.ctypes.data_as(C.c_void_p), _ptr(cov), _ptr(mean), _ptr(out), _ptr(z),
              _stream_ptr(stream))
    xs, zs, off = [], [], 0
    for d, m in zip(dims, samples):
        k = int(d) * int(m)
        xs.append(out[off:off + k].view(int(m), int(d)).t())
        if return_z:
            zs.append(z[off:off + k].view(int(m), int(d)).t())
        off += k
    return (xs, zs) if return_z else xs


def wishart_batch(urng, scales, df, samples, stream=None, inverse=False):
    """Batched WishartGen<float>(df_i, scale_i).generate(engine_i, samples_i)
    (InvWishartGen with inverse=True); each result is (d_i, d_i * samples_i)
    column-major like the reference."""
    torch = _torch()
    n = len(scales)
    dims = np.array([np.asarray(c).shape[0] for c in scales], dtype=np.int32)
    df = np.broadcast_to(np.asarray(df, dtype=np.int64), (n,)).copy()
    samples = np.broadcast_to(np.asarray(samples, dtype=np.int64), (n,)).copy()
    scale = _concat_f64(scales)
    total = int((dims.astype(np.int64) ** 2 * samples).sum())
    out = torch.empty(total, dtype=torch.float32, device="cuda")
    _lib.call("erd_invwishart_batch" if inverse else "erd_wishart_batch", urng.handle, n,
              dims.ctypes.data_as(C.c_void_p),
              df.ctypes.data_as(C.c_void_p), samples.ctypes.data_as(C.c_void_p), _ptr(scale),
              _ptr(out), _stream_ptr(stream))
    res, off = [], 0
    for d, m in zip(dims, samples):
        k = int(d) * int(d) * int(m)
        res.append(out[off:off + k].view(int(d) * int(m), int(d)).t())
        off += k
    return res


def invwishart_batch(urng, scales, df, samples, stream=None):
    """Batched InvWishartGen<float>(df_i, scale_i).generate(engine_i, samples_i)."""
    return wishart_batch(urng, scales, df, samples, stream, inverse=True)


def dirichlet_batch(urng, alphas, samples, stream=None):
    """Batched DirichletGen<float>(alpha_i).generate(engine_i, samples_i);
    each result is (K_i, samples_i)."""
    torch = _torch()
    n = len(alphas)
    dims = np.array([len(a) for a in alphas], dtype=np.int32)
    samples = np.broadcast_to(np.asarray(samples, dtype=np.int64), (n,)).copy()
    alpha = np.ascontiguousarray(np.concatenate([np.asarray(a, np.float32) for a in alphas]))
    total = int((dims.astype(np.int64) * samples).sum())
    out = torch.empty(total, dtype=torch.float32, device="cuda")
    _lib.call("erd_dirichlet_batch", urng.handle, n, dims.ctypes.data_as(C.c_void_p),
              samples.ctypes.data_as(C.c_void_p), alpha.ctypes.data_as(C.c_void_p), _ptr(out),
              _stream_ptr(stream))
    res, off = [], 0
    for k, m in zip(dims, samples):
        c = int(k) * int(m)
        res.append(out[off:off + c].view(int(m), int(k)).t())
        off += c
    return res


class MvFactors:
    """FP64 Cholesky factors of a batch of covariances (or Wishart scales),
    computed once on the device like the generators' constructors
    (MvNormal.h:20-38, 86-98, 196-215) and reused by every generate()."""

    def __init__(self, mats, stream=None, inverse=False, dims=None):
        if dims is None:
            n = len(mats)
            self.dims = np.array([np.asarray(c).shape[0] for c in mats], dtype=np.int32)
            a = _concat_f64(mats)
        else:
            # mats: one flat float64 CUDA tensor, the d_i x d_i matrices back
            # to back, each column-major (no host copy)
            self.dims = np.asarray(dims, dtype=np.int32).ravel().copy()
            n = int(self.dims.size)
            need = int((self.dims.astype(np.int64) ** 2).sum())
            if (not hasattr(mats, "is_cuda") or not mats.is_cuda or mats.dtype != _torch().float64
                    or not mats.is_contiguous() or mats.numel() != need):
                raise ValueError(f"flat covariances: a contiguous float64 CUDA tensor of {need} elements")
            a = mats
        self.info = np.zeros(n, dtype=np.int32)
        h = C.c_void_p()
        _lib.call("erd_mv_factors_create_inverse" if inverse else "erd_mv_factors_create", C.byref(h), n,
                  self.dims.ctypes.data_as(C.c_void_p), _ptr(a), self.info.ctypes.data_as(C.c_void_p),
                  _stream_ptr(stream))
        self._h = h

    @property
    def handle(self):
        return self._h

    def __len__(self):
        return int(self.dims.size)

    def __del__(self):
        h = getattr(self, "_h", None)
        lib = getattr(_lib, "_lib", None) if _lib is not None else None  # None at interpreter exit
        if h is not None and lib is not None:
            lib.erd_mv_factors_destroy(h)
            self._h = None


class MvNormalBatch:
    """Many MvNormalGen<float>(mean_i, cov_i) at once; generate() runs
    instance i on engine i of `urng`, reusing the factors."""

    def __init__(self, covs, means=None, stream=None, dims=None):
        self.factors = MvFactors(covs, stream, dims=dims)
        self.mean = None
        if means is not None:
            self.mean = _torch().from_numpy(
                np.concatenate([np.asarray(m, np.float32).ravel() for m in means])).cuda()

    def generate(self, urng, samples, return_z=False, out=None, stream=None, flat=False):
        """Per-instance (d_i, samples_i) views, or with flat=True the flat
        output buffer (instances back to back, each column-major)."""
        torch = _torch()
        dims = self.factors.dims
        n = dims.size
        samples = np.broadcast_to(np.asarray(samples, dtype=np.int64), (n,)).copy()
        total = int((dims.astype(np.int64) * samples).sum())
        if out is None:
            out = torch.empty(total, dtype=torch.float32, device="cuda")
        z = torch.empty(total, dtype=torch.float32, device="cuda") if return_z else None
        _lib.call("erd_mvnormal_generate", urng.handle, self.factors.handle, samples.ctypes.data_as(C.c_void_p),
                  _ptr(self.mean), _ptr(out), _ptr(z), _stream_ptr(stream))
        if flat:
            return (out, z) if return_z else out
        xs, zs, off = [], [], 0
        for d, m in zip(dims, samples):
            k = int(d) * int(m)
            xs.append(out[off:off + k].view(int(m), int(d)).t())
            if return_z:
                zs.append(z[off:off + k].view(int(m), int(d)).t())
            off += k
        return (xs, zs) if return_z else xs


class WishartBatch:
    """Many WishartGen<float>(df_i, scale_i) at once (factors reused);
    InvWishartGen with inverse=True (factors of the inverse scales)."""

    def __init__(self, scales, df, stream=None, inverse=False, dims=None):
        self.inverse = inverse
        self.factors = MvFactors(scales, stream, inverse=inverse, dims=dims)
        self.df = np.broadcast_to(np.asarray(df, dtype=np.int64), (len(self.factors),)).copy()

    def generate(self, urng, samples, out=None, stream=None, flat=False):
        torch = _torch()
        dims = self.factors.dims
        n = dims.size
        samples = np.broadcast_to(np.asarray(samples, dtype=np.int64), (n,)).copy()
        total = int((dims.astype(np.int64) ** 2 * samples).sum())
        if out is None:
            out = torch.empty(total, dtype=torch.float32, device="cuda")
        _lib.call("erd_invwishart_generate" if self.inverse else "erd_wishart_generate", urng.handle,
                  self.factors.handle, self.df.ctypes.data_as(C.c_void_p),
                  samples.ctypes.data_as(C.c_void_p), _ptr(out), _stream_ptr(stream))
        if flat:
            return out
        res, off = [], 0
        for d, m in zip(dims, samples):
            k = int(d) * int(d) * int(m)
            res.append(out[off:off + k].view(int(d) * int(m), int(d)).t())
            off += k
        return res


class MvNormalGen:
    """Rand::MvNormalGen<float>(mean, cov) on the GPU (MvNormal.h:56-116);
    the covariance is factorised once, at construction."""

    def __init__(self, mean, cov):
        self.mean = np.asarray(mean, dtype=np.float32).ravel()
        self.cov = np.asarray(cov, dtype=np.float64)
        self._batch = MvNormalBatch([self.cov], [self.mean])

    def dims(self):
        return self.mean.size

    def generate(self, urng, samples):
        return self._batch.generate(urng, [samples])[0]


class WishartGen:
    """Rand::WishartGen<float>(df, scale) on the GPU (MvNormal.h:184-302);
    the scale is factorised once, at construction."""

    def __init__(self, df, scale):
        self.df = int(df)
        self.scale = np.asarray(scale, dtype=np.float64)
        self._batch = WishartBatch([self.scale], [self.df])

    def dims(self):
        return self.scale.shape[0]

    def generate(self, urng, samples):
        return self._batch.generate(urng, [samples])[0]


class InvWishartGen(WishartGen):
    """Rand::InvWishartGen<float>(df, scale) on the GPU (MvNormal.h:340-466);
    the inverse scale is formed and factorised once, at construction."""

    def __init__(self, df, scale):
        self.df = int(df)
        self.scale = np.asarray(scale, dtype=np.float64)
        self._batch = WishartBatch([self.scale], [self.df], inverse=True)


class DirichletGen:
    """Rand::DirichletGen<float>(alpha) on the GPU (Multinomial.h:159-217)."""

    def __init__(self, alpha):
        self.alpha = np.asarray(alpha, dtype=np.float32).ravel()

    def dims(self):
        return self.alpha.size

    def generate(self, urng, samples):
        return dirichlet_batch(urng, [self.alpha], [samples])[0]


def alias_table(weights):
    """Host table of DiscreteGen<int32_t, int32_t>: ("cdf", cdf) for K < 16,
    else ("alias", prob, alias) (AliasMethod, Discrete.h:78-192)."""
    w = np.ascontiguousarray(np.asarray(weights, dtype=np.float64))
    K = w.size
    if K < 16:
        cdf = np.zeros(K, dtype=np.uint32)
        _lib.call("erd_alias_table", w.ctypes.data_as(C.c_void_p), K, None, None,
                  cdf.ctypes.data_as(C.c_void_p))
        return ("cdf", cdf)
    size = 1 << int(np.ceil(np.log2(K)))
    prob = np.zeros(size, dtype=np.int32)
    alias = np.zeros(size, dtype=np.int32)
    _lib.call("erd_alias_table", w.ctypes.data_as(C.c_void_p), K, prob.ctypes.data_as(C.c_void_p),
              alias.ctypes.data_as(C.c_void_p), None)
    return ("alias", prob, alias)


def discrete(rows, cols, urng, weights, out=None, stream=None, precision="int32"):
    """Rand::discrete<MatrixXi>(rows, cols, urng, weights) (int32 precision);
    precision="float" is Rand::discreteF (DiscreteGen<int32_t, float>) and
    precision="double" Rand::discreteD (DiscreteGen<int32_t, double>)."""
    torch = _torch()
    w = np.ascontiguousarray(np.asarray(weights, dtype=np.float64))
    per = int(rows) * int(cols)
    buf = _alloc(urng, per, torch.int32, out)
    name = {"int32": "erd_fill_discrete_i32", "float": "erd_fill_discrete_f_i32",
            "double": "erd_fill_discrete_d_i32"}[precision]
    _lib.call(name, urng.handle, _ptr(buf), per, w.ctypes.data_as(C.c_void_p), w.size, _stream_ptr(stream))
    return _shape(buf, urng, rows, cols) if out is None else buf


def multinomial_batch(urng, trials, weights, samples, stream=None):
    """Batched MultinomialGen<int32_t>(trials_i, weights_i).generate(engine_i,
    samples_i); each result is (K_i, samples_i) int32."""
    torch = _torch()
    n = len(weights)
    dims = np.array([len(w) for w in weights], dtype=np.int64)
    trials = np.broadcast_to(np.asarray(trials, dtype=np.int32), (n,)).copy()
    samples = np.broadcast_to(np.asarray(samples, dtype=np.int64), (n,)).copy()
    w = np.ascontiguousarray(np.concatenate([np.asarray(x, np.float64) for x in weights]))
    total = int((dims * samples).sum())
    out = torch.empty(total, dtype=torch.int32, device="cuda")
    _lib.call("erd_multinomial_batch", urng.handle, n, dims.ctypes.data_as(C.c_void_p),
              trials.ctypes.data_as(C.c_void_p), samples.ctypes.data_as(C.c_void_p),
              w.ctypes.data_as(C.c_void_p), _ptr(out), _stream_ptr(stream))
    res, off = [], 0
    for k, m in zip(dims, samples):
        c = int(k) * int(m)
        res.append(out[off:off + c].view(int(m), int(k)).t())
        off += c
    return res


class MultinomialGen:
    """Rand::MultinomialGen<int32_t>(trials, weights) on the GPU (Multinomial.h:25-134)."""

    def __init__(self, trials, weights):
        self.trials = int(trials)
        self.weights = np.asarray(weights, dtype=np.float64).ravel()

    def dims(self):
        return self.weights.size

    def generate(self, urng, samples):
        return multinomial_batch(urng, [self.trials], [self.weights], [samples])[0]
