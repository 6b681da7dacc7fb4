"""Python binding of the serving runtime (erd_server_* in include/eigenrand_amd.h).

A `Server` owns per-client Vmt19937_64 engines resident on the device and a
launcher/completer pair of native threads that batch queued requests into
segment-program launches as streams and HBM budget free up. Engine g of client
c is seeded `splitmix64(c ^ seed_salt) + 4 g` (its four MT19937-64 lanes take
that value .. +3). A request uses `spread(...)[0]` consecutive engines of its
client starting at a per-client cursor that advances (modulo the client's
engine count) in submission order, so a client's results depend only on its
own request sequence.
"""
import ctypes as C

import numpy as np

from . import _lib

MVNORMAL = 100
MULTINOMIAL = 101
ENGINE_WORK = 65536.0  # normal-equivalent samples per engine


class ServerConfig(C.Structure):
    _fields_ = [("max_clients", C.c_int64), ("engines_per_client", C.c_int32),
                ("max_inflight", C.c_int32), ("hbm_budget_bytes", C.c_int64),
                ("max_batch_requests", C.c_int64), ("seed_salt", C.c_uint64),
                ("max_batch_elements", C.c_int64), ("engine_work", C.c_int64),
                ("reject_cap", C.c_int32), ("lockstep_cap", C.c_int32)]


class ServeRequest(C.Structure):
    _fields_ = [("client", C.c_uint64), ("kind", C.c_int32), ("dim", C.c_int32),
                ("count", C.c_int64), ("a", C.c_double), ("b", C.c_double),
                ("host_param", C.c_void_p), ("host_mean", C.c_void_p), ("model", C.c_uint64)]


class ServerStats(C.Structure):
    _fields_ = [("completed", C.c_int64), ("elements", C.c_int64), ("batches", C.c_int64),
                ("p50_ms", C.c_double), ("p99_ms", C.c_double), ("max_ms", C.c_double),
                ("span_s", C.c_double), ("budget_in_use", C.c_int64), ("prep_s", C.c_double)]


def splitmix64(x):
    m = (1 << 64) - 1
    x = (x + 0x9E3779B97F4A7C15) & m
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & m
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & m
    return x ^ (x >> 31)


def client_seeds(client, engines, seed_salt=0):
    """Per-engine seeds of a client's engines."""
    base = splitmix64((int(client) ^ int(seed_salt)) & ((1 << 64) - 1))
    return [(base + 4 * g) & ((1 << 64) - 1) for g in range(engines)]


def chunk_for(kind, a=0.0, b=0.0, dim=0, engine_work=0):
    """Outputs (samples for MvNormal/multinomial) one engine takes; mirrors
    chunk_for in serve.hip (engine_work 0: the default)."""
    if isinstance(kind, str):
        kind = _lib.DIST[kind]
    cost = 1.0
    if kind == 8:        # Student-t
        cost = 2.0
    elif kind == 9:      # gamma
        if a == 1.0:
            cost = 1.0
        elif a == float(int(a)) and a < 20:
            cost = a / 2
        else:
            cost = 5.0 if a < 1 else 6.0
    elif kind == 10:     # Poisson
        cost = 10.0 + 6.0 * a if a < 12 else 7.0
    elif kind == 11:     # binomial
        m = a * min(b, 1 - b)
        cost = 1.0 + 1.2 * a if a < 25 else (13.0 if m < 1 else 12.0)
    elif kind == MVNORMAL:
        cost = float(dim)
    elif kind == MULTINOMIAL:
        cost = 1.0 + a / 8 + dim / 8 if a < max(30.0 * (dim - 1), 100.0) else 12.0 * dim
    work = float(engine_work) if engine_work > 0 else ENGINE_WORK
    return max(1, int(work / max(cost, 1e-3)))


def spread(kind, count, engines_per_client, width, a=0.0, b=0.0, dim=0, engine_work=0):
    """(engines, per-engine outputs or samples) a request is spread over."""
    chunk = max(chunk_for(kind, a, b, dim, engine_work), width)
    e = int(min(engines_per_client, -(-count // chunk)))
    per = -(-count // e)
    per = -(-per // width) * width
    return e, per


_ELEM = {0: (np.uint64, 4), 2: (np.float64, 4), 10: (np.int32, 4), 11: (np.int32, 4)}


def kind_layout(kind):
    """(numpy dtype, packet width) of a scalar-parameter request kind."""
    if kind in _ELEM:
        return _ELEM[kind]
    if kind == MULTINOMIAL:
        return np.int32, 8
    return np.float32, 8


class Server:
    def __init__(self, max_clients=64, engines_per_client=256, max_inflight=8,
                 hbm_budget_bytes=64 << 30, max_batch_requests=256, seed_salt=0,
                 max_batch_elements=0, engine_work=0, reject_cap=0, lockstep_cap=0):
        self.cfg = ServerConfig(max_clients, engines_per_client, max_inflight, hbm_budget_bytes,
                                max_batch_requests, seed_salt & ((1 << 64) - 1), max_batch_elements,
                                engine_work, reject_cap, lockstep_cap)
        h = C.c_void_p()
        _lib.call("erd_server_create", C.byref(h), C.byref(self.cfg))
        self._h = h

    def close(self):
        lib = getattr(_lib, "_lib", None) if _lib is not None else None  # None at interpreter exit
        if getattr(self, "_h", None) is not None and lib is not None:
            lib.erd_server_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def register_mvnormal(self, client, cov, mean=None):
        """Registers a client's MvNormal model; returns its id for submit(model=...)."""
        c = np.asfortranarray(np.asarray(cov, dtype=np.float64))
        m = None if mean is None else np.ascontiguousarray(np.asarray(mean, dtype=np.float32))
        mid = C.c_uint64()
        _lib.call("erd_server_register_mvnormal", self._h, C.c_uint64(int(client) & ((1 << 64) - 1)),
                  int(c.shape[0]), c.ctypes.data_as(C.c_void_p),
                  None if m is None else m.ctypes.data_as(C.c_void_p), C.byref(mid))
        return mid.value

    def submit(self, client, kind, count, a=0.0, b=0.0, cov=None, mean=None, weights=None, dim=0, model=0):
        """Queue one request; returns its ticket. `kind` is a DIST name or
        number (scalar-parameter kinds), MVNORMAL (a registered `model`, or
        cov d x d, optional mean) or MULTINOMIAL (weights K, trials a)."""
        if isinstance(kind, str):
            kind = _lib.DIST[kind]
        req = ServeRequest(int(client) & ((1 << 64) - 1), int(kind), int(dim), int(count),
                           float(a), float(b), None, None, int(model))
        keep = []
        if kind == MVNORMAL and model:
            if mean is not None:
                m = np.ascontiguousarray(np.asarray(mean, dtype=np.float32))
                keep.append(m)
                req.host_mean = m.ctypes.data
        elif kind == MVNORMAL:
            c = np.asfortranarray(np.asarray(cov, dtype=np.float64))
            req.dim = c.shape[0]
            keep.append(c)
            req.host_param = c.ctypes.data
            if mean is not None:
                m = np.ascontiguousarray(np.asarray(mean, dtype=np.float32))
                keep.append(m)
                req.host_mean = m.ctypes.data
        elif kind == MULTINOMIAL:
            w = np.ascontiguousarray(np.asarray(weights, dtype=np.float64))
            req.dim = w.size
            keep.append(w)
            req.host_param = w.ctypes.data
        t = C.c_uint64()
        _lib.call("erd_server_submit", self._h, C.byref(req), C.byref(t))
        return t.value

    def wait(self, ticket, timeout_s=-1.0):
        """Blocks until the request is done; returns (device pointer, elements)."""
        p, n = C.c_void_p(), C.c_int64()
        _lib.call("erd_server_wait", self._h, C.c_uint64(ticket), float(timeout_s), C.byref(p),
                  C.byref(n))
        return p.value, n.value

    def fetch(self, ticket, dtype):
        """Waits and copies the request's output to a host numpy array."""
        p, n = self.wait(ticket)
        out = np.empty(n, dtype=dtype)
        _lib.call("erd_memcpy_d2h", out.ctypes.data_as(C.c_void_p), C.c_void_p(p), out.nbytes, None)
        return out

    def release(self, ticket):
        _lib.call("erd_server_release", self._h, C.c_uint64(ticket))

    def stats(self):
        st = ServerStats()
        _lib.call("erd_server_stats", self._h, C.byref(st))
        return {f: getattr(st, f) for f, _ in ServerStats._fields_}

    def reset_stats(self):
        _lib.call("erd_server_reset_stats", self._h)

    def client_engines(self, client):
        """The client's engines as a non-owning Vmt19937_64 view."""
        from . import Vmt19937_64
        h = C.c_void_p()
        _lib.call("erd_server_client_engines", self._h, C.c_uint64(int(client) & ((1 << 64) - 1)),
                  C.byref(h))
        return Vmt19937_64(_handle=h, _parent=self)
