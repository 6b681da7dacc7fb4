"""erd_run_requests_noalloc against erd_run_requests, through the C ABI: two
calls in a row on the same engines (the second with fewer requests), outputs
and engine records bit-identical; a request count above the scratch limit is
rejected before anything is enqueued."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

SPECS = [  # (engine, dist, count, a, b)
    (0, "normal", 24, 1.0, 2.0), (1, "gamma", 16, 0.5, 1.0), (0, "poisson", 12, 20.0, 0),
    (2, "binomial", 8, 30, 0.4), (1, "uniform", 40, -1.0, 1.0), (0, "gamma", 8, 3.7, 1.0),
    (3, "std_normal", 4096, 0, 0), (1, "poisson", 20, 3.0, 0), (2, "binomial", 4, 10, 0.7),
    (0, "exponential", 8, 2.0, 0), (1, "cauchy", 8, 0.0, 1.0), (2, "student_t", 16, 5.0, 0),
    (3, "gamma", 4096, 4.0, 0.5), (2, "discard", 7, 0, 0), (2, "u64", 8, 0, 0),
    (1, "binomial", 8, 1000, 0.001), (0, "uniform_f64", 8, 0.0, 1.0), (3, "poisson", 2048, 7.5, 0),
    (4, "normal_vop", 64, 0, 0), (4, "binomial_vop", 32, 0, 0), (5, "gamma_vop", 64, 0, 0),
    (5, "poisson_vop", 32, 0, 0), (3, "lognormal", 16, 0.0, 1.0), (5, "weibull", 16, 1.7, 2.5),
]
DTYPES = {"poisson": torch.int32, "binomial": torch.int32, "binomial_vop": torch.int32,
          "poisson_vop": torch.int32, "u64": torch.uint64, "uniform_f64": torch.float64}


def build(specs, rng):
    """(ctypes request array, outputs, parameter tensors kept alive)."""
    from eigenrand_amd import _lib
    arr = (_lib.Request * len(specs))()
    outs, keep = [], []
    for i, (eng, dist, count, a, b) in enumerate(specs):
        o = None if dist == "discard" else torch.empty(count, dtype=DTYPES.get(dist, torch.float32),
                                                       device="cuda")
        in0 = in1 = None
        if dist == "normal_vop":
            in0 = torch.from_numpy(rng.uniform(-3, 3, count).astype(np.float32)).cuda()
            in1 = torch.from_numpy(rng.uniform(0.5, 2, count).astype(np.float32)).cuda()
        elif dist == "binomial_vop":
            in0 = torch.from_numpy(rng.integers(1, 500, count).astype(np.int32)).cuda()
            in1 = torch.from_numpy(rng.uniform(0, 1, count).astype(np.float32)).cuda()
        elif dist == "gamma_vop":
            in0 = torch.from_numpy(np.exp(rng.uniform(-2, 3, count)).astype(np.float32)).cuda()
            in1 = torch.ones(count, dtype=torch.float32, device="cuda")
        elif dist == "poisson_vop":
            in0 = torch.from_numpy(np.exp(rng.uniform(-4, 8, count)).astype(np.float32)).cuda()
        arr[i].dist = _lib.DIST[dist]
        arr[i].engine = eng
        arr[i].count = count
        arr[i].out = o.data_ptr() if o is not None else None
        arr[i].in0 = in0.data_ptr() if in0 is not None else None
        arr[i].in1 = in1.data_ptr() if in1 is not None else None
        arr[i].a = float(a)
        arr[i].b = float(b)
        outs.append(o)
        keep += [in0, in1]
    return arr, outs, keep


def run(name, urng, arr):
    from eigenrand_amd import _lib
    rc = getattr(_lib.load(), name)(urng.handle, C.cast(arr, C.c_void_p), len(arr), None)
    assert rc == 0, (name, rc)


def test_noalloc_requests_equal_run_requests_twice():
    import eigenrand_amd as er
    seeds = [100, 200, 300, 400, 500, 600]
    ua = er.Vmt19937_64(seeds=seeds)   # erd_run_requests
    ub = er.Vmt19937_64(seeds=seeds)   # erd_run_requests_noalloc
    for call, specs in enumerate((SPECS, SPECS[3:9] + SPECS[18:20])):
        rng_a = np.random.default_rng(40 + call)
        rng_b = np.random.default_rng(40 + call)
        arr_a, outs_a, keep_a = build(specs, rng_a)
        arr_b, outs_b, keep_b = build(specs, rng_b)
        run("erd_run_requests", ua, arr_a)
        run("erd_run_requests_noalloc", ub, arr_b)
        torch.cuda.synchronize()
        for (eng, dist, *_), oa, ob in zip(specs, outs_a, outs_b):
            if oa is None:
                continue
            assert torch.equal(oa.view(torch.uint8), ob.view(torch.uint8)), (call, eng, dist)
        np.testing.assert_array_equal(ua.snapshot(), ub.snapshot())
        assert ua.status() == 0 and ub.status() == 0


def test_noalloc_rejects_more_requests_than_its_scratch_holds():
    import eigenrand_amd as er
    from eigenrand_amd import _lib
    lib = _lib.load()
    u = er.Vmt19937_64(seeds=[1])
    n = 4097  # ERD_NOALLOC_MAX_REQUESTS + 1
    arr = (_lib.Request * n)()
    snap = u.snapshot()
    rc = lib.erd_run_requests_noalloc(u.handle, C.cast(arr, C.c_void_p), n, None)
    assert rc == 1  # ERD_ERR_INVALID
    np.testing.assert_array_equal(u.snapshot(), snap)
    # the largest allowed count runs (discards of one packet on one engine)
    arr = (_lib.Request * 4096)()
    for i in range(4096):
        arr[i].dist = _lib.DIST["discard"]
        arr[i].engine = 0
        arr[i].count = 1
    assert lib.erd_run_requests_noalloc(u.handle, C.cast(arr, C.c_void_p), 4096, None) == 0
    ref = er.Vmt19937_64(seeds=[1])
    ref.discard(4096)
    np.testing.assert_array_equal(u.snapshot(), ref.snapshot())
