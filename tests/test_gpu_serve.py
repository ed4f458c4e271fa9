"""Serving (dbx_server): continuous batching of a stream of models of uneven size.

Every request's result must be the same alone and in the mix: histograms bit-exact against the
same model run alone through dbx_intensity, I(q) within 1e-9 relative of that run (both use the
matrix-core pass 2; only the FP64 order of the split sums differs) and within 1e-6 of the FP64
oracle.  Re-weighting a resident request replays a captured graph; each replay is checked
against the direct computation with the same form factors."""
import threading
import time

import numpy as np
import pytest

import cudebye_amd as cd
from cudebye_amd import synth
from oracle import oracle as O
from tests.test_gpu_pass1 import _oracle_hist

pytestmark = pytest.mark.gpu

MIX_TOL = 1e-9
ORACLE_TOL = 1e-6


@pytest.fixture(scope="module")
def ctx():
    c = cd.Context(0)
    yield c
    c.close()


def _gp(p):
    return cd.GroupParams.make(p["occ"], p["beq"], p["a"], p["b"], p["c"])


def _oracle_I(m, H, params=None):
    params = params or m.params

    class Pm:
        pass
    Pm.group_n = m.group_n
    Pm.occ = np.array([p["occ"] for p in params])
    Pm.beq = np.array([p["beq"] for p in params])
    Pm.a = np.array([p["a"] for p in params])
    Pm.b = np.array([p["b"] for p in params])
    Pm.c = np.array([p["c"] for p in params])
    return O.intensity_from_histograms(H, Pm, m.q, m.rstep, first_n=m.first_n)


def _models(n, seed, nqs=(64, 128, 96)):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        na = int(np.exp(rng.uniform(np.log(100), np.log(3000))))
        S = int(rng.integers(1, 5))
        out.append(synth.small_random(na, S, seed * 1000 + i, spread=float(rng.uniform(4.0, 15.0)),
                                      nq=int(nqs[i % len(nqs)])))
    return out


def _alone(ctx, m):
    mod = ctx.model(m.x, m.y, m.z, m.group_n)
    I, H = ctx.intensity(mod, [_gp(p) for p in m.params], m.rstep, m.n_bins, m.q, first_n=m.first_n,
                         with_hist=True)
    mod.close()
    return I, H


def _submit(srv, m, keep=True, n_bins=None, params=None):
    return srv.submit(m.x, m.y, m.z, m.group_n, [_gp(p) for p in (params or m.params)], m.rstep,
                      n_bins or m.n_bins, m.q, first_n=m.first_n, keep=keep)


def _rel(a, b):
    return np.max(np.abs(a - b) / np.abs(b))


def _finish(srv, t, timeout=60.0, allow=()):
    """srv.wait(t) after polling until the request has ended, so that a request that never
    ends fails the test instead of blocking it."""
    deadline = time.time() + timeout
    while srv.poll(t) not in (cd.DBX_REQ_DONE, cd.DBX_REQ_FAILED):
        assert time.time() < deadline, f"request {t} did not finish within {timeout} s"
        time.sleep(0.005)
    return srv.wait(t, allow=allow)


def test_serve_results_same_alone_and_in_the_mix(ctx):
    models = _models(16, 3)
    alone = [_alone(ctx, m) for m in models]
    with cd.Server(0) as srv:
        tickets = [None] * len(models)

        def late():
            # arrivals while the first ones are on the GPU
            for i in range(8, 16):
                time.sleep(0.002)
                tickets[i] = _submit(srv, models[i])

        th = threading.Thread(target=late)
        for i in range(8):
            tickets[i] = _submit(srv, models[i])
        th.start()
        th.join()
        worst = 0.0
        for i, m in enumerate(models):
            I, lat, rc = _finish(srv, tickets[i])
            assert rc == cd.DBX_OK and lat > 0
            H = srv.histogram(tickets[i])
            assert np.array_equal(H, alone[i][1]), f"request {i} histogram"
            assert _rel(I, alone[i][0]) < MIX_TOL, f"request {i}"
            if i % 4 == 0:
                ref, _ = _oracle_hist(m)
                assert np.array_equal(H, ref)
                worst = max(worst, _rel(I, _oracle_I(m, ref)))
            srv.release(tickets[i])
        st = srv.stats()
        print(f"steps {st['steps']} max step {st['max_step_requests']} worst vs oracle {worst:.2e}")
        assert worst < ORACLE_TOL
        assert st["completed"] == len(models) and st["resident_bytes"] == 0


def test_serve_reweight_replays_match_direct(ctx):
    m = _models(1, 7, nqs=(200,))[0]
    rng = np.random.default_rng(11)
    sets = [[synth.cromer_mann(rng) for _ in m.params] for _ in range(3)]
    with cd.Server(0) as srv:
        t = _submit(srv, m)
        I0, _, _ = _finish(srv, t)
        H = srv.histogram(t)
        mod = ctx.model(m.x, m.y, m.z, m.group_n)
        for rep, k in enumerate([0, 1, 0, 2, 1]):
            Ir = srv.reweight(t, [_gp(p) for p in sets[k]])
            # eager work on the device between replays
            Id = ctx.intensity(mod, [_gp(p) for p in sets[k]], m.rstep, m.n_bins, m.q, first_n=m.first_n)
            assert _rel(Ir, Id) < MIX_TOL, f"replay {rep}"
            assert _rel(Ir, _oracle_I(m, H, sets[k])) < ORACLE_TOL
        assert _rel(srv.reweight(t, [_gp(p) for p in m.params]), I0) < 1e-12
        srv.release(t)
        with pytest.raises(cd.DebyeError):
            srv.reweight(t, [_gp(p) for p in m.params])



def test_serve_release_while_another_thread_reweights(ctx):
    """dbx_release on one thread while another reweights the same request: each reweight
    either returns the request's pattern or, once the request is gone, fails with an argument
    or state error (never reads freed memory)."""
    m = _models(1, 9, nqs=(128,))[0]
    params = [_gp(p) for p in m.params]
    with cd.Server(0) as srv:
        t = _submit(srv, m)
        I0, _, _ = _finish(srv, t)
        ok, errs = [], []
        go = threading.Event()

        def loop():
            go.wait()
            for _ in range(300):
                try:
                    ok.append(_rel(srv.reweight(t, params), I0))
                except cd.DebyeError as e:
                    errs.append(e.code)
                    break

        th = threading.Thread(target=loop)
        th.start()
        go.set()
        time.sleep(0.003)
        srv.release(t)
        th.join()
        assert all(r < 1e-12 for r in ok)
        assert all(c in (cd.DBX_ERR_ARG, cd.DBX_ERR_STATE) for c in errs)
        assert srv.stats()["resident_bytes"] == 0
        # a released ticket is unknown to every call that takes one
        _, _, rc = srv.wait(t, allow=(cd.DBX_ERR_ARG,))
        assert rc == cd.DBX_ERR_ARG
        with pytest.raises(cd.DebyeError):
            srv.histogram(t)
        with pytest.raises(cd.DebyeError):
            srv.release(t)

def test_serve_admission_waits_for_budget():
    # five requests of nearly the same size; a budget that holds two of them resident
    models = [synth.small_random(1500, 2, 500 + i, spread=10.0, nq=64) for i in range(5)]
    sizes = []
    with cd.Server(0) as srv:
        for m in models:
            t = _submit(srv, m)
            _finish(srv, t)
            sizes.append(srv.stats()["resident_bytes"])
            srv.release(t)
    assert min(sizes) > 0 and max(sizes) < 1.2 * min(sizes)
    budget = int(2.5 * max(sizes))
    with cd.Server(0, hbm_budget=budget) as srv:
        ts = [_submit(srv, m) for m in models]
        _finish(srv, ts[0])
        _finish(srv, ts[1])
        time.sleep(0.3)
        states = [srv.poll(t) for t in ts]
        assert states[:2] == [cd.DBX_REQ_DONE] * 2
        assert states[2:] == [cd.DBX_REQ_QUEUED] * 3, states
        assert srv.stats()["resident_bytes"] <= budget
        # each release admits the next waiting request
        for k in range(2, 5):
            srv.release(ts[k - 2])
            _, _, rc = _finish(srv, ts[k])
            assert rc == cd.DBX_OK
            assert srv.stats()["resident_bytes"] <= budget
        assert srv.stats()["completed"] == len(models)


def test_serve_request_over_budget_refused():
    m = _models(1, 17)[0]
    with cd.Server(0, hbm_budget=1 << 16) as srv:
        with pytest.raises(cd.DebyeError) as e:
            _submit(srv, m, n_bins=1 << 20)
        assert e.value.code == cd.DBX_ERR_BUDGET


def test_serve_bad_request_fails_alone(ctx):
    models = _models(4, 19, nqs=(64,))
    short = O.channels(5.0, models[1].rstep)
    alone = [_alone(ctx, m) for m in models]
    with cd.Server(0) as srv:
        ts = [_submit(srv, m, n_bins=short if i == 1 else None) for i, m in enumerate(models)]
        for i, t in enumerate(ts):
            I, _, rc = _finish(srv, t, allow=(cd.DBX_ERR_RANGE,))
            if i == 1:
                assert rc == cd.DBX_ERR_RANGE and srv.poll(t) == cd.DBX_REQ_FAILED
            else:
                assert rc == cd.DBX_OK
                assert np.array_equal(srv.histogram(t), alone[i][1])
                assert _rel(I, alone[i][0]) < MIX_TOL
            srv.release(t)


def test_serve_poisson_stream_small(ctx):
    rng = np.random.default_rng(23)
    models = _models(24, 29)
    with cd.Server(0) as srv:
        ts = []
        for m in models:
            time.sleep(rng.exponential(0.002))
            ts.append(_submit(srv, m, keep=False))
        lats = []
        for t in ts:
            I, lat, rc = _finish(srv, t)
            assert rc == cd.DBX_OK and np.all(np.isfinite(I))
            lats.append(lat)
            srv.release(t)
        st = srv.stats()
        print(f"p50 {np.percentile(lats, 50):.2f} ms p99 {np.percentile(lats, 99):.2f} ms steps {st['steps']}")
        assert st["completed"] == len(models) and st["resident_bytes"] == 0


# Requests of the mixed serving stream (synth.serve_request): every q grid (2048, 4096, 8192
# points), 1-4 species, one of 121k atoms and two below 1k.
C5_REQUESTS = (7, 6, 13, 1, 2, 3, 15, 18, 0, 57)


@pytest.fixture(scope="module")
def c5_reference(ctx):
    """The C5 requests, each run alone through dbx_intensity, and the oracle's histograms and
    FP64 intensities of every one."""
    models = [synth.serve_request(i) for i in C5_REQUESTS]
    assert {len(m.q) for m in models} == {2048, 4096, 8192}
    assert {len(m.group_n) for m in models} == {1, 2, 3, 4}
    assert max(m.n_atoms for m in models) >= 100000 and min(m.n_atoms for m in models) <= 1000
    alone = [_alone(ctx, m) for m in models]
    oracle = []
    for m, (_, H) in zip(models, alone):
        print(f"C5 request {m.name}: {m.n_atoms} atoms, {len(m.group_n)} species, {len(m.q)} Q", flush=True)
        ref, ov = _oracle_hist(m)
        assert ov == 0
        oracle.append((ref, _oracle_I(m, ref)))
    return models, alone, oracle


def _check_c5(srv, models, tickets, alone, oracle):
    worst = 0.0
    for m, t, (Ia, Ha), (Ho, Io) in zip(models, tickets, alone, oracle):
        _finish(srv, t, timeout=120.0)
        I, lat, rc, bound = srv.wait(t, with_err=True)
        assert rc == cd.DBX_OK and lat > 0
        H = srv.histogram(t)
        assert np.array_equal(H, Ha), m.name
        assert np.array_equal(H, Ho), m.name
        assert _rel(I, Ia) < MIX_TOL, m.name
        # each I within its reported bound of the oracle (1e-14: the oracle's own prefactors)
        assert bound.max() <= ORACLE_TOL and np.all(np.abs(I - Io) / np.abs(Io) <= bound + 1e-14), m.name
        worst = max(worst, _rel(I, Io))
        srv.release(t)
    assert worst < ORACLE_TOL, worst
    return worst


def test_serve_c5_requests_in_one_step(c5_reference):
    """The C5 requests submitted at once form one step (a batching window covers their host
    preparation): one pass-1 launch over all of them and one contraction per q grid, three grids
    in the step.  Each result: histograms bit-exact against the request alone and the oracle, I
    within 1e-9 of the request alone and 1e-6 of the oracle's FP64 sums."""
    models, alone, oracle = c5_reference
    with cd.Server(0, batch_window_ms=5000.0) as srv:
        ts = [_submit(srv, m) for m in models]
        worst = _check_c5(srv, models, ts, alone, oracle)
        st = srv.stats()
        print(f"steps {st['steps']} requests/step {st['max_step_requests']} q grids/step "
              f"{st['max_step_qgrids']} worst vs oracle {worst:.2e}")
        assert st["steps"] == 1 and st["max_step_requests"] == len(models) and st["max_step_qgrids"] == 3
        assert st["resident_bytes"] == 0


def test_serve_c5_requests_arriving_over_time(c5_reference):
    """The same requests arriving one by one from another thread, largest first, into a server
    that starts steps at once: whatever steps they fall into, each result is the same as alone."""
    models, alone, oracle = c5_reference
    order = sorted(range(len(models)), key=lambda i: -models[i].n_atoms)
    with cd.Server(0) as srv:
        ts = [None] * len(models)

        def feed():
            for i in order:
                ts[i] = _submit(srv, models[i])
                time.sleep(0.003)

        th = threading.Thread(target=feed)
        th.start()
        th.join()
        _check_c5(srv, models, ts, alone, oracle)
        print(f"steps {srv.stats()['steps']}")


def test_serve_kept_requests_over_budget_do_not_block_wait():
    """More DBX_REQ_KEEP requests than the budget holds, waited on in order without releasing:
    once nothing runs that could free memory, dbx_wait on a request still waiting for memory
    returns DBX_ERR_BUDGET instead of blocking; releasing a kept request admits it and a later
    wait returns its result."""
    models = [synth.small_random(1500, 2, 600 + i, spread=10.0, nq=64) for i in range(4)]
    with cd.Server(0) as srv:
        t = _submit(srv, models[0])
        _finish(srv, t)
        size = srv.stats()["resident_bytes"]
        srv.release(t)
    with cd.Server(0, hbm_budget=int(2.5 * size)) as srv:
        ts = [_submit(srv, m) for m in models]
        rcs = []
        for t in ts:
            _, _, rc = srv.wait(t, allow=(cd.DBX_ERR_BUDGET,))
            rcs.append(rc)
        assert rcs[:2] == [cd.DBX_OK] * 2 and rcs[2:] == [cd.DBX_ERR_BUDGET] * 2, rcs
        assert [srv.poll(t) for t in ts[2:]] == [cd.DBX_REQ_QUEUED] * 2
        srv.release(ts[0])
        srv.release(ts[1])
        for t in ts[2:]:
            _, _, rc = _finish(srv, t)
            assert rc == cd.DBX_OK
            srv.release(t)
        assert srv.stats()["resident_bytes"] == 0


def test_serve_close_while_another_thread_waits():
    """dbx_server_close while another thread is blocked in dbx_wait: the waiter returns (the
    result, or DBX_ERR_STATE once the server stops) and the close completes."""
    m = synth.serve_request(6)
    srv = cd.Server(0)
    t = _submit(srv, m)
    out = []

    def waiter():
        _, _, rc = srv.wait(t, allow=(cd.DBX_ERR_STATE,))
        out.append(rc)

    th = threading.Thread(target=waiter)
    th.start()
    time.sleep(0.001)
    srv.close()
    th.join(timeout=60)
    assert not th.is_alive() and out and out[0] in (cd.DBX_OK, cd.DBX_ERR_STATE)


def test_serve_error_target_recomputes_in_fp64(ctx):
    """A server whose error target lies below the matrix-core bounds: every request's partials
    are recomputed with FP64 arithmetic (the step's requests of one q grid in one pass 2), giving
    the FP64 result of the request alone to within both results' bounds, with a bound of the FP64
    kind (its rounding term grows with the compacted bins of the whole pass 2); a reweight of a
    kept request reports the bound for its new form factors."""
    models = _models(5, 41, nqs=(96, 160))
    ctx.set_pass2_mode(cd.DBX_PASS2_FP64)
    try:
        ref = []
        for m in models:
            mod = ctx.model(m.x, m.y, m.z, m.group_n)
            ref.append(ctx.intensity(mod, [_gp(p) for p in m.params], m.rstep, m.n_bins, m.q, first_n=m.first_n,
                                     with_err=True))
            mod.close()
    finally:
        ctx.set_pass2_mode(cd.DBX_PASS2_MFMA)
    with cd.Server(0, rel_err_target=1e-15) as srv:
        ts = [_submit(srv, m) for m in models]
        for m, t, (I64, b64) in zip(models, ts, ref):
            _finish(srv, t)
            I, _, rc, bound = srv.wait(t, with_err=True)
            assert rc == cd.DBX_OK
            assert np.all(np.abs(I - I64) / np.abs(I64) <= bound + b64)
            assert np.all(bound <= 8 * b64), (bound / b64).max()
        Ir, br = srv.reweight(ts[0], [_gp(p) for p in models[0].params], with_err=True)
        I0, _, _, b0 = srv.wait(ts[0], with_err=True)
        assert _rel(Ir, I0) < 1e-13 and np.allclose(br, b0, rtol=1e-9, atol=0)
        for t in ts:
            srv.release(t)


def test_serve_mixed_bin_widths(ctx):
    """Requests of three bin widths (1e-3, 5e-4 and 2e-3 A) submitted interleaved: a step takes
    the head's bin width only, so they run in separate steps; every request's histograms are
    bit-exact against the model alone at its own width and its I(q) within the mix tolerance."""
    base = _models(12, 17)
    widths = [1e-3, 5e-4, 2e-3]
    models = []
    for i, m in enumerate(base):
        w = widths[i % 3]
        models.append(synth.small_random(int(m.group_n.sum()), len(m.group_n), 17000 + i,
                                         spread=float(np.ptp(m.x)) / 2 + 1.0, rstep=w, nq=len(m.q)))
    alone = [_alone(ctx, m) for m in models]
    with cd.Server(0) as srv:
        tickets = [_submit(srv, m) for m in models]
        for i, m in enumerate(models):
            I, _, rc = _finish(srv, tickets[i])
            assert rc == cd.DBX_OK
            H = srv.histogram(tickets[i])
            assert np.array_equal(H, alone[i][1]), f"request {i} (bin {m.rstep})"
            assert _rel(I, alone[i][0]) < MIX_TOL, f"request {i}"
            srv.release(tickets[i])
        st = srv.stats()
        assert st["completed"] == len(models) and st["steps"] >= 3 and st["resident_bytes"] == 0


def test_serve_fp64_mode(ctx):
    """A server configured for FP64 pass 2 (the reference's arithmetic): every request's
    histograms bit-exact against the model alone and its I(q) within 1e-11 of the FP64 mode's
    result alone (the same sums over another split of the bins), and within both bounds."""
    models = _models(6, 53, nqs=(80, 120))
    ctx.set_pass2_mode(cd.DBX_PASS2_FP64)
    try:
        ref = []
        for m in models:
            mod = ctx.model(m.x, m.y, m.z, m.group_n)
            ref.append(ctx.intensity(mod, [_gp(p) for p in m.params], m.rstep, m.n_bins, m.q, first_n=m.first_n,
                                     with_hist=True, with_err=True))
            mod.close()
    finally:
        ctx.set_pass2_mode(cd.DBX_PASS2_MFMA)
    with cd.Server(0, pass2_mode=cd.DBX_PASS2_FP64) as srv:
        ts = [_submit(srv, m) for m in models]
        for t, (I64, H64, b64) in zip(ts, ref):
            _finish(srv, t)
            I, _, rc, bound = srv.wait(t, with_err=True)
            assert rc == cd.DBX_OK
            assert np.array_equal(srv.histogram(t), H64)
            assert _rel(I, I64) < 1e-11
            assert np.all(np.abs(I - I64) / np.abs(I64) <= bound + b64)
            srv.release(t)
