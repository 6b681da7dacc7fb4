"""Iteration caps: a sampler that gives up does so through its kernel's own
bound (every kernel exits normally), and the failure reaches the caller on
both paths: the engine set's status (plain fills) and the request status of
the server, which then returns no output. The caps are forced low from the
host; nothing here makes a kernel run away.

Also: a model's first request publishes its factor planes only once they are
complete, so a second request of the same model on another stream reads the
finished factor.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from test_gpu_serve import Client, _spd, _srv  # noqa: E402


def _cap_cases(er):
    """Fills whose lockstep loops check the cap at every step and need more
    than two iterations somewhere with near certainty (the windowed Knuth
    fills and the gamma samplers finish their packets within one step, so a
    cap of 1 cannot bite there)."""
    n = 4096
    mean = torch.full((64 * n,), 5.0, device="cuda")
    trials = torch.full((64 * n,), 400, dtype=torch.int32, device="cuda")
    p_knuth = torch.full((64 * n,), 0.002, device="cuda")   # mean 0.8: Knuth lockstep
    p_tan = torch.full((64 * n,), 0.3, device="cuda")       # tangent, 16 iterations a step
    return [
        ("poisson VoP, Knuth", lambda u: er.poisson_vop(u, mean)),
        ("binomial VoP, Knuth", lambda u: er.binomial_vop(u, trials, p_knuth)),
        ("binomial VoP, tangent", lambda u: er.binomial_vop(u, trials, p_tan)),
    ]


def test_fills_give_up_at_the_lockstep_cap_and_report_it(er):
    for name, fill in _cap_cases(er):
        u = er.Vmt19937_64(11, n=64)
        u.set_iteration_caps(lockstep_cap=1)  # at most two iterations per packet
        fill(u)
        torch.cuda.synchronize()
        with pytest.raises(er.ErdError) as ei:
            u.check()
        assert ei.value.status == 4, name
        u.check()  # the bits were cleared by the read
        # the default caps: the same fill completes and reports nothing
        v = er.Vmt19937_64(11, n=64)
        fill(v)
        v.check()


def test_views_copy_caps_and_bad_caps_are_rejected(er):
    u = er.Vmt19937_64(3, n=8)
    u.set_iteration_caps(lockstep_cap=1)
    v = u.view(0, 8)
    er.poisson_vop(v, torch.full((8 * 4096,), 5.0, device="cuda"))
    with pytest.raises(er.ErdError):
        v.check()
    with pytest.raises(er.InvalidArgument):
        u.set_iteration_caps(reject_cap=-1)


def test_server_request_at_the_cap_fails_without_output():
    import eigenrand_amd as er
    from eigenrand_amd import _lib
    from eigenrand_amd.serving import MULTINOMIAL, spread
    E = 8
    # trials >= 30 (K - 1): the conditional-binomial rows; the first row's
    # binomial(1000, 3e-4) runs the Knuth lockstep loop
    w = np.array([1.0, 1e-4, 1e-4, 1e-4])
    with _srv(lockstep_cap=1, max_batch_requests=1) as s:
        t_bad = s.submit(1, MULTINOMIAL, 8_000, a=1000, weights=w)
        p, n = C.c_void_p(), C.c_int64()
        rc = _lib.load().erd_server_wait(s._h, C.c_uint64(t_bad), -1.0, C.byref(p), C.byref(n))
        assert rc == 4  # ERD_ERR_ITER_CAP
        assert not p.value  # no output block is handed out
        s.release(t_bad)
        # a sampler that needs no iteration cap, for another client, succeeds
        t_ok = s.submit(2, "std_normal", 50_000)
        got = s.fetch(t_ok, np.float32)
        s.release(t_ok)
        # the failed client's next request reseeds its engines
        t_again = s.submit(1, "std_normal", 20_000)
        again = s.fetch(t_again, np.float32)
    e, per = spread("std_normal", 50_000, E, 8)
    want = []
    for v in Client(2, E).parts(e):
        want.append(er.normal(per, 1, v).reshape(-1).cpu().numpy())
    np.testing.assert_array_equal(got, np.concatenate(want))
    cl = Client(1, E)
    e_bad, _ = spread(MULTINOMIAL, 8_000, E, 8, a=1000, dim=len(w))
    cl.parts(e_bad)  # the failed request still moved the client's cursor
    e2, per2 = spread("std_normal", 20_000, E, 8)
    want = []
    for v in cl.parts(e2):
        want.append(er.normal(per2, 1, v).reshape(-1).cpu().numpy())
    np.testing.assert_array_equal(again, np.concatenate(want))


def test_second_request_of_a_model_on_another_stream_reads_the_finished_factor():
    import eigenrand_amd as er
    from eigenrand_amd.serving import MVNORMAL, spread
    rng = np.random.default_rng(5)
    E, d = 8, 200
    cov = _spd(d, rng)
    with _srv(max_batch_requests=1, max_inflight=4) as s:
        mid = s.register_mvnormal(9, cov)
        ta = s.submit(9, MVNORMAL, 1_000, model=mid)    # factorises the model (engines 0..3)
        tb = s.submit(9, MVNORMAL, 500, model=mid)      # engines 4, 5: next batch and stream
        ga = s.fetch(ta, np.float32)
        gb = s.fetch(tb, np.float32)
    cl = Client(9, E)
    for got, n in ((ga, 1_000), (gb, 500)):
        e, p = spread(MVNORMAL, n, E, 8, dim=d)
        want = []
        for v in cl.parts(e):
            xs = er.mvnormal_batch(v, [cov] * len(v), [p] * len(v))
            want += [x.t().contiguous().cpu().numpy().ravel() for x in xs]
        np.testing.assert_array_equal(got, np.concatenate(want))
