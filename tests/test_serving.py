"""Scheduling of the shot server (simwave_amd/serving.py) with test doubles:
admission against the HBM budget and the concurrency cap, arrival order,
oversized requests, failures that stay local, latency statistics.  The GPU
behaviour (results identical alone and in the mix) is in
test_gpu_serving.py."""
import threading
import time

import pytest

from simwave_amd.serving import ShotServer, percentile


class FakeRequest:
    def __init__(self, ident, nbytes, steps=20, fail=False):
        self.ident = ident
        self.device_bytes = nbytes
        self.steps = steps
        self.fail = fail


class Monitor:
    def __init__(self):
        self.lock = threading.Lock()
        self.used = 0
        self.active = 0
        self.peak_used = 0
        self.peak_active = 0
        self.order = []


class FakeShot:
    def __init__(self, req, mon):
        self.req, self.mon = req, mon
        self.steps = req.steps
        self.updates_per_shot = 1000 * req.steps
        self.left = req.steps
        self.planned = False

    def plan(self, stream, flags=0):
        with self.mon.lock:
            self.mon.used += self.req.device_bytes
            self.mon.active += 1
            self.mon.peak_used = max(self.mon.peak_used, self.mon.used)
            self.mon.peak_active = max(self.mon.peak_active, self.mon.active)
            self.mon.order.append(self.req.ident)
        self.planned = True

    def run_steps(self, n, stream):
        if self.req.fail:
            raise RuntimeError("plan run failed: injected")
        time.sleep(0.0005)
        self.left = max(0, self.left - n)
        return self.left

    def results(self, stream=None):
        return ("u", self.req.ident), ("rec", self.req.ident)

    def release_plan(self):
        if self.planned:
            with self.mon.lock:
                self.mon.used -= self.req.device_bytes
                self.mon.active -= 1
            self.planned = False


class FakeStream:
    def synchronize(self):
        pass


def make_server(mon, **kw):
    return ShotServer(shot_factory=lambda req: FakeShot(req, mon),
                      stream_factory=FakeStream, **kw)


def test_budget_and_cap_are_respected():
    mon = Monitor()
    reqs = [FakeRequest(i, nbytes=(i % 3 + 1) * 10, steps=30) for i in range(40)]
    with make_server(mon, hbm_budget=45, max_active=3, chunk_steps=4) as srv:
        futs = [srv.submit(r) for r in reqs]
    for r, f in zip(reqs, futs):
        assert f.result(timeout=30) == (("u", r.ident), ("rec", r.ident))
    assert mon.peak_used <= 45
    assert mon.peak_active <= 3
    # admitted in arrival order (the workers may start in either order)
    assert srv.admissions == list(range(len(reqs)))
    assert sorted(mon.order) == [r.ident for r in reqs]
    s = srv.stats.summary()
    assert s["requests"] == 40 and s["max_active"] <= 3
    assert s["updates"] == sum(1000 * r.steps for r in reqs)
    assert s["latency_p99_s"] >= s["latency_p50_s"] > 0


def test_concurrency_happens():
    mon = Monitor()
    reqs = [FakeRequest(i, nbytes=1, steps=200) for i in range(8)]
    with make_server(mon, hbm_budget=100, max_active=8, chunk_steps=2) as srv:
        futs = [srv.submit(r) for r in reqs]
    [f.result(timeout=30) for f in futs]
    assert mon.peak_active > 1


def test_oversized_and_failing_requests_fail_alone():
    mon = Monitor()
    reqs = [FakeRequest(0, 10), FakeRequest(1, 1000), FakeRequest(2, 10, fail=True),
            FakeRequest(3, 10)]
    with make_server(mon, hbm_budget=50, max_active=2) as srv:
        futs = [srv.submit(r) for r in reqs]
    assert futs[0].result(timeout=30)[0] == ("u", 0)
    with pytest.raises(MemoryError):
        futs[1].result(timeout=30)
    with pytest.raises(RuntimeError, match="injected"):
        futs[2].result(timeout=30)
    assert futs[3].result(timeout=30)[0] == ("u", 3)
    assert mon.used == 0 and mon.active == 0              # everything released


def test_submit_after_close_raises():
    mon = Monitor()
    srv = make_server(mon)
    srv.close()
    with pytest.raises(RuntimeError):
        srv.submit(FakeRequest(0, 1))


class FakeBatchShot(FakeShot):
    """A shot the server advances through its batch runner."""
    batchable = True
    variant = "fake"
    rounds = []

    @staticmethod
    def run_batched(shots, steps, stream):
        FakeBatchShot.rounds.append(len(shots))
        time.sleep(0.001)
        out = []
        for s in shots:
            s.left = max(0, s.left - steps)
            out.append(s.left)
        return out


def test_batchable_shots_advance_together():
    mon = Monitor()
    FakeBatchShot.rounds = []
    reqs = [FakeRequest(i, nbytes=1, steps=10 + 7 * i) for i in range(12)]
    with ShotServer(shot_factory=lambda req: FakeBatchShot(req, mon),
                    stream_factory=FakeStream, hbm_budget=100, max_active=12,
                    chunk_steps=5) as srv:
        futs = [srv.submit(r) for r in reqs]
        for r, f in zip(reqs, futs):
            assert f.result(timeout=30) == (("u", r.ident), ("rec", r.ident))
    assert max(FakeBatchShot.rounds) > 1          # several shots per batched call
    s = srv.stats.summary()
    assert s["batched"] == len(reqs)
    assert mon.used == 0 and mon.active == 0


def test_percentile_matches_numpy_linear():
    vals = [5.0, 1.0, 3.0, 2.0, 4.0]
    assert percentile(vals, 50) == 3.0
    assert abs(percentile(vals, 99) - 4.96) < 1e-12


def test_results_dropped_on_request():
    """keep_result=False: the shot still runs and its results are fetched
    (timed in the stats), but the Future holds None."""
    mon = Monitor()
    reqs = [FakeRequest(i, nbytes=1, steps=10) for i in range(6)]
    with make_server(mon, hbm_budget=100, max_active=4, chunk_steps=4) as srv:
        futs = [srv.submit(r, keep_result=(r.ident % 2 == 0)) for r in reqs]
    for r, f in zip(reqs, futs):
        got = f.result(timeout=30)
        assert got == ((("u", r.ident), ("rec", r.ident)) if r.ident % 2 == 0 else None)
    s = srv.stats.summary()
    assert s["requests"] == 6 and s["updates"] == sum(1000 * r.steps for r in reqs)


@pytest.mark.parametrize("ahead", [0, 3])
def test_callable_requests_are_built_once_in_order(ahead):
    """Requests given as callables are built once each (ahead of admission
    on the builder threads with build_ahead > 0, on the scheduler thread
    with 0), admitted in submission order, and a build that raises fails
    only its own request."""
    mon = Monitor()
    calls, threads, lock = [], set(), threading.Lock()

    def builder(i):
        def build():
            with lock:
                calls.append(i)
                threads.add(threading.current_thread().name)
            time.sleep(0.002)
            if i == 5:
                raise ValueError("bad model")
            return FakeRequest(i, 10)
        return build

    with make_server(mon, hbm_budget=25, max_active=2, build_ahead=ahead) as srv:
        futs = [srv.submit(builder(i)) for i in range(12)]
    for i, f in enumerate(futs):
        if i == 5:
            with pytest.raises(ValueError, match="bad model"):
                f.result(timeout=30)
        else:
            assert f.result(timeout=30)[0] == ("u", i)
    assert sorted(calls) == list(range(12))               # each built once
    assert mon.order == [i for i in range(12) if i != 5]  # arrival order kept
    assert srv.admissions == [i for i in range(12) if i != 5]
    if ahead:
        assert all(t.startswith("shot-build") for t in threads), threads
    else:
        assert threads == {"shot-scheduler"}
    assert mon.used == 0 and mon.active == 0
