"""Serving runtime on the GPU: results equal single fills on the client's
engines, are the same alone and in a batched mix, and HBM admission holds
requests back until outputs are released."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _srv(**kw):
    from eigenrand_amd.serving import Server
    cfg = dict(max_clients=16, engines_per_client=8, max_inflight=4, hbm_budget_bytes=1 << 30,
               max_batch_requests=64, seed_salt=0x5EED)
    cfg.update(kw)
    return Server(**cfg)


def _direct_engines(client, E, salt=0x5EED):
    import eigenrand_amd as er
    from eigenrand_amd.serving import client_seeds
    return er.Vmt19937_64(seeds=client_seeds(client, E, salt))


def _spd(d, rng):
    a = rng.normal(0, 1 / np.sqrt(d), (d, d))
    return a @ a.T + d * np.eye(d)


class Client:
    """Replays a client's requests on plain engine sets seeded like the
    server's, following its engine cursor."""

    def __init__(self, client, E, salt=0x5EED):
        self.eng = _direct_engines(client, E, salt)
        self.E = E
        self.cursor = 0

    def parts(self, e):
        """Contiguous views covering engines cursor .. cursor+e-1 (mod E)."""
        c, E = self.cursor, self.E
        self.cursor = (c + e) % E
        if c + e <= E:
            return [self.eng.view(c, e)]
        return [self.eng.view(c, E - c), self.eng.view(0, c + e - E)]


def test_serve_univariate_equals_direct_fills():
    import eigenrand_amd as er
    from eigenrand_amd.serving import spread
    E = 8
    reqs = [("std_normal", 200_000, 0.0, 0.0, 8, torch.float32, np.float32,
             lambda v, p, o: er.normal(p, 1, v, out=o)),
            ("gamma", 70_001, 3.7, 2.0, 8, torch.float32, np.float32,
             lambda v, p, o: er.gamma(p, 1, v, 3.7, 2.0, out=o)),
            ("poisson", 1000, 16.0, 0.0, 4, torch.int32, np.int32,
             lambda v, p, o: er.poisson(p, 1, v, 16.0, out=o)),
            ("poisson", 100_000, 3.5, 0.0, 4, torch.int32, np.int32,
             lambda v, p, o: er.poisson(p, 1, v, 3.5, out=o)),
            ("binomial", 300_000, 100, 0.75, 4, torch.int32, np.int32,
             lambda v, p, o: er.binomial(p, 1, v, 100, 0.75, out=o))]
    with _srv() as s:
        tickets = [s.submit(7, k, n, a=a, b=b) for k, n, a, b, *_ in reqs]
        got = [s.fetch(t, r[6]) for t, r in zip(tickets, reqs)]
    cl = Client(7, E)
    engines_used = []
    for (kind, n, a, b, width, tdt, ndt, fill), g in zip(reqs, got):
        e, p = spread(kind, n, E, width, a, b)
        engines_used.append(e)
        outs = []
        for v in cl.parts(e):
            o = torch.empty(len(v) * p, dtype=tdt, device="cuda")
            fill(v, p, o)
            outs.append(o.cpu().numpy())
        np.testing.assert_array_equal(g, np.concatenate(outs), err_msg=kind)
    assert engines_used[1] > 1 and engines_used[-1] == E  # wrapped and overlapping ranges


def test_serve_mvnormal_and_multinomial_equal_batch_calls():
    import eigenrand_amd as er
    from eigenrand_amd.serving import MVNORMAL, MULTINOMIAL, spread
    rng = np.random.default_rng(11)
    E = 8
    d = 24
    cov = _spd(d, rng)
    mean = rng.normal(size=d).astype(np.float32)
    w_small = rng.uniform(0.1, 2.0, 12)     # trials below 30*(K-1): alias/CDF path
    w_rows = rng.uniform(0.1, 2.0, 5)       # trials >= 30*(K-1): binomial rows
    with _srv() as s:
        tm = s.submit(3, MVNORMAL, 9000, cov=cov, mean=mean)
        ts = s.submit(4, MULTINOMIAL, 1003, a=50, weights=w_small)
        tr = s.submit(5, MULTINOMIAL, 20000, a=1000, weights=w_rows)
        tr2 = s.submit(5, MULTINOMIAL, 64, a=1000, weights=w_rows)
        gm = s.fetch(tm, np.float32)
        gs = s.fetch(ts, np.int32)
        gr = s.fetch(tr, np.int32)
        gr2 = s.fetch(tr2, np.int32)
    em, pm = spread(MVNORMAL, 9000, E, 8, dim=d)
    assert em > 1
    want = []
    for v in Client(3, E).parts(em):
        xs = er.mvnormal_batch(v, [cov] * len(v), [pm] * len(v), means=[mean] * len(v))
        want += [x.t().contiguous().cpu().numpy().ravel() for x in xs]
    np.testing.assert_array_equal(gm, np.concatenate(want))
    c4, c5 = Client(4, E), Client(5, E)
    for got, cl, w, T, n in ((gs, c4, w_small, 50, 1003), (gr, c5, w_rows, 1000, 20000),
                             (gr2, c5, w_rows, 1000, 64)):
        e, p = spread(MULTINOMIAL, n, E, 8, a=T, dim=len(w))
        want = []
        for v in cl.parts(e):
            ms = er.multinomial_batch(v, [T] * len(v), [w] * len(v), [p] * len(v))
            want += [m.t().contiguous().cpu().numpy().ravel() for m in ms]
        np.testing.assert_array_equal(got, np.concatenate(want))
        assert (got.reshape(-1, len(w)).sum(axis=1) == T).all()


def _mix(rng, n_req, clients):
    from eigenrand_amd.serving import MVNORMAL, MULTINOMIAL
    covs = {c: _spd(8 + 4 * c, rng) for c in clients}
    reqs = []
    for i in range(n_req):
        c = int(rng.choice(clients))
        u = rng.uniform()
        if u < 0.6:
            kind = ["std_normal", "gamma", "binomial", "poisson", "uniform", "normal"][i % 6]
            n = int(np.exp(rng.uniform(np.log(1e3), np.log(4e5))))
            a, b = {"std_normal": (0, 0), "gamma": (float(rng.uniform(0.3, 8)), 1.5),
                    "binomial": (int(rng.integers(1, 300)), float(rng.uniform(0.05, 0.95))),
                    "poisson": (float(np.exp(rng.uniform(-3, 6))), 0),
                    "uniform": (-1.0, 3.0), "normal": (2.0, 0.5)}[kind]
            reqs.append(dict(client=c, kind=kind, count=n, a=a, b=b))
        elif u < 0.9:
            d = covs[c].shape[0]
            reqs.append(dict(client=c, kind=MVNORMAL, count=int(rng.integers(1, 3000)), cov=covs[c],
                             mean=np.full(d, 0.5, np.float32)))
        else:
            K = int(rng.integers(3, 40))
            reqs.append(dict(client=c, kind=MULTINOMIAL, count=int(rng.integers(1, 2000)),
                             a=int(rng.choice([7, 5000])), weights=rng.uniform(0.01, 1, K)))
    return reqs


def _dtype(r):
    from eigenrand_amd.serving import MULTINOMIAL
    return np.int32 if r["kind"] in ("binomial", "poisson", MULTINOMIAL) else np.float32


def test_serve_results_same_alone_and_in_mix():
    rng = np.random.default_rng(2024)
    reqs = _mix(rng, 48, [1, 2, 3, 9])
    with _srv(max_batch_requests=64, max_inflight=4) as s:
        tickets = [s.submit(**r) for r in reqs]
        mixed = [s.fetch(t, _dtype(r)) for t, r in zip(tickets, reqs)]
        st = s.stats()
    assert st["completed"] == len(reqs)
    assert st["batches"] < len(reqs)  # requests were actually batched
    with _srv(max_batch_requests=1, max_inflight=1) as s:
        alone = []
        for r in reqs:
            t = s.submit(**r)
            alone.append(s.fetch(t, _dtype(r)))
            s.release(t)
        assert s.stats()["batches"] == len(reqs)
    for i, (x, y) in enumerate(zip(mixed, alone)):
        np.testing.assert_array_equal(x, y, err_msg=f"request {i}: {reqs[i]['kind']}")


def test_serve_hbm_admission_waits_for_release():
    from eigenrand_amd._lib import ErdError
    n = 1 << 20  # 4 MiB of float output per request
    with _srv(hbm_budget_bytes=6 << 20) as s:
        t1 = s.submit(1, "std_normal", n)
        t2 = s.submit(2, "std_normal", n)
        s.wait(t1)
        with pytest.raises(ErdError):
            s.wait(t2, timeout_s=0.5)  # held back: 8 MiB would exceed 6 MiB
        assert s.stats()["budget_in_use"] == 4 << 20
        s.release(t1)
        s.wait(t2, timeout_s=30)
        s.release(t2)
        assert s.stats()["budget_in_use"] == 0


def test_serve_rejects_bad_requests():
    from eigenrand_amd._lib import ErdError
    from eigenrand_amd.serving import MULTINOMIAL
    with _srv(max_clients=1) as s:
        with pytest.raises(ErdError):
            s.submit(1, "normal_vop", 100)           # per-element parameters are not served
        with pytest.raises(ErdError):
            s.submit(1, "gamma", 100, a=-1.0, b=1.0)
        with pytest.raises(ErdError):
            s.submit(1, MULTINOMIAL, 10, a=5, weights=[0.0, 0.0])
        with pytest.raises(ErdError):
            s.submit(1, "std_normal", 1 << 30)        # output above the HBM budget
        s.wait(s.submit(1, "std_normal", 64))
        with pytest.raises(ErdError):
            s.submit(2, "std_normal", 64)             # engine slots exhausted


def test_serve_registered_model_equals_inline_covariance():
    from eigenrand_amd.serving import MVNORMAL
    rng = np.random.default_rng(5)
    cov = _spd(40, rng)
    mean = rng.normal(size=40).astype(np.float32)
    with _srv() as s:
        t = s.submit(2, MVNORMAL, 3000, cov=cov, mean=mean)
        inline = s.fetch(t, np.float32)
    with _srv() as s:
        mid = s.register_mvnormal(2, cov, mean)
        t = s.submit(2, MVNORMAL, 3000, model=mid)
        reg = s.fetch(t, np.float32)
        with pytest.raises(Exception):
            s.submit(3, MVNORMAL, 10, model=mid)  # another client's model
    np.testing.assert_array_equal(inline, reg)
