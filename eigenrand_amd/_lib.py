"""ctypes binding of the C ABI in include/eigenrand_amd.h.

The shared library is built in-tree (eigenrand_amd/liberd_amd.so). There is no
fallback: if the library is missing or cannot be loaded, every entry point
raises, so a product call can never silently run on the CPU.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liberd_amd.so")

ERD_OK = 0
STATUS_NAMES = {
    0: "ERD_OK", 1: "ERD_ERR_INVALID", 2: "ERD_ERR_NONFINITE", 3: "ERD_ERR_NOT_PSD",
    4: "ERD_ERR_ITER_CAP", 5: "ERD_ERR_OOM", 6: "ERD_ERR_HIP", 7: "ERD_ERR_UNSUPPORTED",
}


class ErdError(RuntimeError):
    """A C ABI call returned a status other than ERD_OK."""

    def __init__(self, fn, status):
        self.status = status
        super().__init__(f"{fn} failed: {STATUS_NAMES.get(status, status)}")


class InvalidArgument(ErdError, ValueError):
    pass


P = C.c_void_p
I64 = C.c_int64
U64 = C.c_uint64
F32 = C.c_float
F64 = C.c_double
INT = C.c_int

# name -> argtypes (all return int status unless listed in _RESTYPES)
_SIGS = {
    "erd_status_string": [INT],
    "erd_engine_state_bytes": [],
    "erd_engines_create": [C.POINTER(P), P, I64, P],
    "erd_engines_create_range": [C.POINTER(P), U64, U64, I64, P],
    "erd_engines_view": [C.POINTER(P), P, I64, I64],
    "erd_engines_destroy": [P],
    "erd_engines_count": [P],
    "erd_engines_layout": [P],
    "erd_engines_device_ptr": [P],
    "erd_engines_snapshot": [P, P, P],
    "erd_engines_restore": [P, P, P],
    "erd_engines_copy": [P, P, P],
    "erd_engines_discard": [P, I64, P],
    "erd_engines_status": [P, C.POINTER(C.c_uint32), P],
    "erd_engines_set_iteration_caps": [P, C.c_int32, C.c_int32],
    "erd_engines_set_layout": [P, C.c_int],
    "erd_fill_u64": [P, P, I64, P],
    "erd_engines_create_p8": [C.POINTER(P), P, I64, P],
    "erd_p8_fill_u64": [P, P, I64, P],
    "erd_p8_fill_u32": [P, P, I64, P],
    "erd_p8_fill_uniform_f32": [P, P, I64, P],
    "erd_p8_fill_dist": [P, C.c_int32, P, I64, F64, F64, P],
    "erd_fill_uniform_f32": [P, P, I64, F32, F32, P],
    "erd_fill_uniform_f64": [P, P, I64, F64, F64, P],
    "erd_fill_std_normal_f32": [P, P, I64, P],
    "erd_fill_std_normal_f64": [P, P, I64, P],
    "erd_fill_normal_f64": [P, P, I64, F64, F64, P],
    "erd_fill_exponential_f64": [P, P, I64, F64, P],
    "erd_fill_cauchy_f64": [P, P, I64, F64, F64, P],
    "erd_fill_normal_f32": [P, P, I64, F32, F32, P],
    "erd_fill_normal_vop_vs_f32": [P, P, I64, P, F32, P],
    "erd_fill_discrete_d_i32": [P, P, I64, P, I64, P],
    "erd_fill_normal_vop_sv_f32": [P, P, I64, F32, P, P],
    "erd_fill_binomial_vop_vs_i32": [P, P, I64, P, F32, P],
    "erd_fill_binomial_vop_f32": [P, P, I64, P, P, P],
    "erd_fill_binomial_vop_vs_f32": [P, P, I64, P, F32, P],
    "erd_fill_binomial_vop_sv_i32": [P, P, I64, INT, P, P],
    "erd_fill_std_normal_f32_fast": [P, P, I64, P],
    "erd_fill_normal_f32_fast": [P, P, I64, F32, F32, P],
    "erd_store_ceiling_f32": [P, I64, I64, P],
    "erd_fill_normal_vop_f32": [P, P, I64, P, P, P],
    "erd_fill_exponential_f32": [P, P, I64, F32, P],
    "erd_fill_cauchy_f32": [P, P, I64, F32, F32, P],
    "erd_fill_lognormal_f32": [P, P, I64, F32, F32, P],
    "erd_fill_weibull_f32": [P, P, I64, F32, F32, P],
    "erd_fill_extreme_value_f32": [P, P, I64, F32, F32, P],
    "erd_fill_studentt_f32": [P, P, I64, F32, P],
    "erd_fill_gamma_f32": [P, P, I64, F32, F32, P],
    "erd_fill_chisquared_f32": [P, P, I64, F32, P],
    "erd_fill_poisson_i32": [P, P, I64, F64, P],
    "erd_fill_binomial_i32": [P, P, I64, C.c_int32, F64, P],
    "erd_fill_binomial_vop_i32": [P, P, I64, P, P, P],
    "erd_fill_gamma_vop_f32": [P, P, I64, P, P, P],
    "erd_fill_poisson_vop_i32": [P, P, I64, P, P],
    "erd_fill_bernoulli_f32": [P, P, I64, F64, P],
    "erd_fill_bernoulli_vop_f32": [P, P, I64, P, P],
    "erd_fill_balanced_f32": [P, P, I64, F32, F32, P],
    "erd_fill_balanced_vop_f32": [P, P, I64, P, P, P],
    "erd_fill_uniform_vop_f32": [P, P, I64, P, P, P],
    "erd_fill_exponential_vop_f32": [P, P, I64, P, P],
    "erd_fill_cauchy_vop_f32": [P, P, I64, P, P, P],
    "erd_fill_lognormal_vop_f32": [P, P, I64, P, P, P],
    "erd_fill_studentt_vop_f32": [P, P, I64, P, P],
    "erd_fill_weibull_vop_f32": [P, P, I64, P, P, P],
    "erd_fill_extreme_value_vop_f32": [P, P, I64, P, P, P],
    "erd_fill_bernoulli_i32": [P, P, I64, F64, P],
    "erd_fill_geometric_i32": [P, P, I64, F64, P],
    "erd_fill_randbits_i32": [P, P, I64, P],
    "erd_fill_uniform_int_i32": [P, P, I64, C.c_int32, C.c_int32, P],
    "erd_fill_discrete_f_i32": [P, P, I64, P, I64, P],
    "erd_fill_beta_f32": [P, P, I64, F32, F32, P],
    "erd_fill_negative_binomial_i32": [P, P, I64, C.c_int32, F64, P],
    "erd_fill_fisher_f_f32": [P, P, I64, F32, F32, P],
    "erd_generator_create": [C.POINTER(P), C.c_int32, F64, F64, I64, P],
    "erd_generator_destroy": [P],
    "erd_generator_reset": [P, P],
    "erd_generator_fill": [P, P, P, I64, P],
    "erd_run_requests": [P, P, I64, P],
    "erd_run_requests_noalloc": [P, P, I64, P],
    "erd_cholesky_f64_batch": [I64, P, P, P, P],
    "erd_mvnormal_batch": [P, I64, P, P, P, P, P, P, P],
    "erd_wishart_batch": [P, I64, P, P, P, P, P, P],
    "erd_mv_factors_create": [C.POINTER(P), I64, P, P, P, P],
    "erd_mv_factors_destroy": [P],
    "erd_mv_factors_count": [P],
    "erd_mvnormal_generate": [P, P, P, P, P, P, P],
    "erd_wishart_generate": [P, P, P, P, P, P],
    "erd_mv_factors_create_inverse": [C.POINTER(P), I64, P, P, P, P],
    "erd_invwishart_generate": [P, P, P, P, P, P],
    "erd_invwishart_batch": [P, I64, P, P, P, P, P, P],
    "erd_dirichlet_batch": [P, I64, P, P, P, P, P],
    "erd_alias_table": [P, I64, P, P, P],
    "erd_fill_discrete_i32": [P, P, I64, P, I64, P],
    "erd_multinomial_batch": [P, I64, P, P, P, P, P, P],
    "erd_server_create": [C.POINTER(P), P],
    "erd_server_destroy": [P],
    "erd_server_submit": [P, P, C.POINTER(U64)],
    "erd_server_register_mvnormal": [P, U64, C.c_int32, P, P, C.POINTER(U64)],
    "erd_server_wait": [P, U64, F64, C.POINTER(P), C.POINTER(I64)],
    "erd_server_release": [P, U64],
    "erd_server_stats": [P, P],
    "erd_server_reset_stats": [P],
    "erd_server_client_engines": [P, U64, C.POINTER(P)],
    "erd_device_alloc": [C.POINTER(P), C.c_size_t],
    "erd_device_free": [P],
    "erd_memcpy_d2h": [P, P, C.c_size_t, P],
    "erd_memcpy_h2d": [P, P, C.c_size_t, P],
    "erd_stream_synchronize": [P],
}
_RESTYPES = {
    "erd_status_string": C.c_char_p,
    "erd_engine_state_bytes": C.c_size_t,
    "erd_engines_count": I64,
    "erd_mv_factors_count": I64,
    "erd_engines_device_ptr": P,
}

_lib = None


def load():
    """Load the native library (once) and declare every signature."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"eigenrand_amd native library not built ({LIB_PATH}); run __graft_entry__.build()")
    lib = C.CDLL(LIB_PATH)
    for name, args in _SIGS.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = _RESTYPES.get(name, INT)
    _lib = lib
    return lib


def call(name, *args):
    """Call a status-returning entry point and raise on failure."""
    rc = getattr(load(), name)(*args)
    if rc != ERD_OK:
        if rc == 1:
            raise InvalidArgument(name, rc)
        raise ErdError(name, rc)
    return rc


def declared_symbols():
    return list(_SIGS)


class Request(C.Structure):
    """erd_request (include/eigenrand_amd.h)."""
    _fields_ = [("dist", C.c_int32), ("engine", C.c_int32), ("count", C.c_int64),
                ("out", C.c_void_p), ("in0", C.c_void_p), ("in1", C.c_void_p),
                ("a", C.c_double), ("b", C.c_double)]


DIST = {
    "u64": 0, "uniform": 1, "uniform_f64": 2, "std_normal": 3, "normal": 4, "normal_vop": 5,
    "exponential": 6, "cauchy": 7, "student_t": 8, "gamma": 9, "poisson": 10, "binomial": 11,
    "binomial_vop": 12, "gamma_vop": 13, "poisson_vop": 14, "discard": 15, "lognormal": 16,
    "weibull": 17, "extreme_value": 18,
}
