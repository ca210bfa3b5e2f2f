"""The streaming service's scheduling (svdimager_amd/serve.py) with a fake device: admission
against the budget, batches of one (image size, cell) key capped at max_batch, oldest request
first, completion bookkeeping and the Poisson driver; and the batching-form snapshot draws
(svdimager_amd/synth.py)."""
import dataclasses

import numpy as np
import pytest

from svdimager_amd import serve, synth


class FakeSnap:
    def __init__(self, nbytes, samples=1000):
        self.nbytes = nbytes
        self.num_samples = samples


class FakeClock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t


class FakeDevice:
    """Each batched call completes `latency` clock units after it is launched."""

    def __init__(self, clock, latency=1.0):
        self.clock = clock
        self.latency = latency
        self.launched = []

    def __call__(self, reqs, key):
        self.launched.append((key, [r.rid for r in reqs]))
        done_at = self.clock() + self.latency
        clock = self.clock

        class H:
            def query(self):
                return clock() >= done_at
        return H()


def make_service(budget, ws=1000, **kw):
    clock = FakeClock()
    dev = FakeDevice(clock)
    svc = serve.SnapshotService(None, budget, clock=clock, launcher=dev, workspace_bytes=lambda: ws,
                                **kw)
    return svc, clock, dev


def req(rid, N=64, cell=1e-4, nbytes=10_000, t=0.0):
    return serve.Request(rid, FakeSnap(nbytes), N, cell, arrival=t)


def drain(svc, clock, limit=1000):
    for _ in range(limit):
        svc.poll()
        if not svc.busy():
            return
        clock.t += 0.5
    raise AssertionError("did not drain")


def test_admission_keeps_within_budget():
    cost = serve.request_bytes(req(0))
    svc, clock, dev = make_service(budget=1000 + int(2.5 * cost), max_batch=8, max_inflight=4)
    for i in range(6):
        svc.submit(req(i, t=0.0))
    seen = []
    for _ in range(100):
        svc.poll()
        assert svc.workspace + svc.held <= svc.budget
        seen.append(svc.held)
        if not svc.busy():
            break
        clock.t += 0.5
    assert svc.peak_held == 2 * cost
    assert [r.rid for r in svc.completed] == list(range(6))
    assert all(b <= 2 for b in svc.batch_sizes)


def test_batches_share_a_key_and_respect_max_batch():
    svc, clock, dev = make_service(budget=10**12, max_batch=3, max_inflight=1)
    keys = [(64, 1e-4), (128, 1e-4), (64, 2e-4)]
    for i in range(12):
        N, cell = keys[i % 3]
        svc.submit(req(i, N=N, cell=cell, t=float(i)))
    drain(svc, clock)
    assert len(svc.completed) == 12
    for key, rids in dev.launched:
        assert 1 <= len(rids) <= 3
        assert all(keys[r % 3] == key for r in rids)
    # The first batch takes the key of the oldest request, with the oldest requests of that key.
    assert dev.launched[0] == ((64, 1e-4), [0, 3, 6])


def test_oldest_key_first_and_inflight_cap():
    svc, clock, dev = make_service(budget=10**12, max_batch=8, max_inflight=2)
    svc.submit(req(0, N=128, t=0.0))
    svc.submit(req(1, N=64, t=1.0))
    svc.submit(req(2, N=256, t=2.0))
    svc.poll()
    assert [k for k, _ in dev.launched] == [(128, 1e-4), (64, 1e-4)]   # two in flight, third waits
    clock.t = 1.0
    svc.poll()
    assert [k for k, _ in dev.launched][-1] == (256, 1e-4)
    drain(svc, clock)
    r = {q.rid: q for q in svc.completed}
    assert r[0].done == r[1].done == 1.0 and r[2].done >= 2.0
    assert all(q.latency == q.done - q.arrival for q in svc.completed)


def test_workspace_growth_past_the_budget_is_an_error():
    ws = [1000]
    clock = FakeClock()
    svc = serve.SnapshotService(None, 1000 + int(1.5 * serve.request_bytes(req(0))), clock=clock,
                                launcher=FakeDevice(clock), workspace_bytes=lambda: ws[0],
                                max_inflight=1)
    svc.submit(req(0))
    svc.submit(req(1))
    svc.poll()                                   # request 0 admitted and launched, 1 waits
    assert [r.rid for r in svc.waiting] == [1]
    ws[0] = svc.budget
    svc.workspace = ws[0]
    clock.t = 5.0
    with pytest.raises(RuntimeError):
        svc.poll()                               # request 0 retired; request 1 can never fit


def test_request_over_budget_is_refused():
    svc, clock, dev = make_service(budget=1000 + 100)
    with pytest.raises(ValueError):
        svc.submit(req(0, nbytes=10_000))


def test_poisson_arrivals():
    counts = []
    for seed in range(40):
        t = serve.poisson_arrivals(200.0, 0.5, np.random.default_rng(seed))
        assert (np.diff(t) > 0).all() and (t >= 0).all() and (t < 0.5).all()
        counts.append(len(t))
    # Mean 100 per draw; the mean of 40 draws has standard deviation sqrt(100 / 40) ~ 1.6.
    assert abs(np.mean(counts) - 100.0) < 8.0
    assert len(serve.poisson_arrivals(0.0, 1.0, np.random.default_rng(0))) == 0


def test_run_poisson_with_fake_device():
    import time
    dev_clock = time.perf_counter
    launched = []

    def launcher(reqs, key):
        launched.append(len(reqs))
        done_at = dev_clock() + 2e-3

        class H:
            def query(self):
                return dev_clock() >= done_at
        return H()

    svc = serve.SnapshotService(None, 10**12, max_batch=4, max_inflight=2, launcher=launcher,
                                workspace_bytes=lambda: 0)
    pool = [(FakeSnap(1000, 500), 64, 1e-4), (FakeSnap(2000, 700), 128, 1e-4)]
    out = serve.run_poisson(svc, pool, rate=2000.0, duration=0.1, seed=3, drain_limit=5.0)
    assert out["count"] == len(out["requests"]) > 100
    lat = out["latency_ms"]
    assert 2.0 <= lat["p50"] <= lat["p99"] <= lat["max"]
    assert max(launched) <= 4 and out["mean_batch"] >= 1.0
    assert all(r.done >= r.launched >= r.admitted >= r.arrival for r in out["requests"])


def test_batching_draws_and_row_removal():
    cfg = dataclasses.replace(synth.SMALL, flag_fraction=0.0)
    draws = [synth.batching_draw(i, 9, cfg, sizes=(64, 128)) for i in range(9)]
    assert draws[0][0] == -4.0 and draws[-1][0] == 4.0 and draws[4][0] == 0.0
    assert all(0.0 <= d[1] < 0.6 and d[2] in (64, 128) for d in draws)
    full = synth.make_snapshot(cfg, ha_hours=draws[2][0], seed_offset=2)
    cut = synth.make_batching_snapshot(2, 9, cfg, sizes=(64, 128))
    R, C = full["num_rows"], full["num_channels"]
    Rk = cut["num_rows"]
    assert Rk == round(R * (1.0 - draws[2][1])) and cut["image_size"] == draws[2][2]
    # Every kept row appears in the same order with all its channels (idx = row + chan * R).
    keep = np.array([int(np.flatnonzero((full["uvw"] == u).all(axis=1))[0]) for u in cut["uvw"]])
    assert (np.diff(keep) > 0).all()
    for k in synth.PER_SAMPLE_FIELDS:
        np.testing.assert_array_equal(cut[k].reshape(C, Rk), full[k].reshape(C, R)[:, keep])
