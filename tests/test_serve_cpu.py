"""Admission and ordering of the serving loop, on the CPU with stand-in engines.

The stand-ins record which handle ran what and when; no GPU is involved.
Checks: every request completes; the resident model sees its updates and
breaker operations in arrival order; areas start first-come first-served and
never hold more than the HBM budget at once; an area larger than the whole
budget is refused instead of blocking the queue; a failing request does not
stop the loop.
"""
import threading
import time

import numpy as np

from state_estimation_with_cuda_amd import serve


class FakeGrid:
    def __init__(self, n):
        self.n = n


class FakeEngine:
    def __init__(self, name, log, delay=0.002, fail_on=None):
        self.name, self.log, self.delay, self.fail_on = name, log, delay, fail_on
        self.lock = threading.Lock()  # a handle is never used from two threads at once

    def set_telemetry_indexed(self, idx, val):
        assert self.lock.acquire(blocking=False)
        try:
            if self.fail_on is not None and int(idx[0]) == self.fail_on:
                raise RuntimeError("bad telemetry")
            self.log.append((self.name, "z", int(idx[0])))
        finally:
            self.lock.release()

    def set_branch_status(self, k, st):
        self.log.append((self.name, "brk", k, st))

    def solve(self, **kw):
        assert self.lock.acquire(blocking=False)
        try:
            time.sleep(self.delay)
            return {"status": 0, "iterations": 3}
        finally:
            self.lock.release()

    def state(self, out=None):
        return np.zeros(2), np.zeros(2)

    def gpu_bytes(self):
        return 0

    def close(self):
        pass


def make_requests(n_upd=60, rate=600.0, areas=(300, 500, 200, 900, 400), seed=0):
    rng = np.random.default_rng(seed)
    t = serve.poisson_arrivals(rate, n_upd / rate, rng)
    reqs = [serve.Request(serve.UPDATE, float(a), (np.array([k], np.int32), np.array([1.0])), rid=k)
            for k, a in enumerate(t)]
    reqs.append(serve.Request(serve.TOGGLE, float(t[len(t) // 2]) + 1e-6, (7, 0), rid=10_000))
    for j, n in enumerate(areas):
        reqs.append(serve.Request(serve.AREA, 0.01 * j, FakeGrid(n), rid=20_000 + j))
    return reqs


def test_all_requests_complete_in_order_and_within_budget():
    log = []
    main = FakeEngine("main", log)
    peak = {"now": 0, "max": 0}
    lock = threading.Lock()

    class AreaEngine(FakeEngine):
        def __init__(self, grid):
            super().__init__(f"area{grid.n}", log, delay=0.01)
            self.bytes = grid.n * 100
            with lock:
                peak["now"] += self.bytes
                peak["max"] = max(peak["max"], peak["now"])

        def close(self):
            with lock:
                peak["now"] -= self.bytes

    budget = 100 * 1000  # room for about two areas at a time
    srv = serve.Server(main, AreaEngine, hbm_budget=budget, area_slots=4, bytes_per_bus=100.0)
    reqs = make_requests()
    done = serve.run(srv, reqs)
    assert len(done) == len(reqs)
    assert all(r.error is None for r in done)
    assert srv.peak_used <= budget and peak["max"] <= budget
    # resident model: updates in arrival order (the breaker between them)
    seq = [e[2] for e in log if e[0] == "main" and e[1] == "z"]
    assert seq == sorted(seq)
    # areas start in arrival order
    areas = sorted((r for r in done if r.kind == serve.AREA), key=lambda r: r.start)
    assert [r.rid for r in areas] == sorted(r.rid for r in areas)
    s = serve.summarize(done)
    n_upd = sum(1 for r in reqs if r.kind == serve.UPDATE)
    assert s["update"]["count"] == n_upd and s["update"]["p99_ms"] is not None


def test_oversized_area_is_refused_and_failures_are_isolated():
    log = []
    main = FakeEngine("main", log, fail_on=5)
    srv = serve.Server(main, lambda g: FakeEngine("a", log), hbm_budget=1000, area_slots=2, bytes_per_bus=10.0)
    reqs = make_requests(n_upd=20, areas=(50, 500, 60))  # 500 buses x 10 B > budget
    done = serve.run(srv, reqs)
    assert len(done) == len(reqs)
    big = [r for r in done if r.kind == serve.AREA and r.payload.n == 500][0]
    assert big.error and "budget" in big.error
    bad = [r for r in done if r.kind == serve.UPDATE and r.rid == 5][0]
    assert bad.error and "bad telemetry" in bad.error
    assert sum(1 for r in done if r.error is None) == len(reqs) - 2


def test_queued_area_whose_estimate_outgrows_the_budget_fails_instead_of_blocking():
    """An area engine that uses more HBM than estimated raises bytes_per_bus; an
    area queued meanwhile whose new estimate exceeds the whole budget must fail
    rather than wait forever (and hold up every area behind it)."""
    log = []
    main = FakeEngine("main", log)
    started = threading.Event()
    release = threading.Event()

    class Overrun(FakeEngine):
        def __init__(self, grid):
            super().__init__(f"area{grid.n}", log, delay=0.0)
            self.n = grid.n
            if grid.n == 100:  # the first area: hold it until the big one is queued
                started.set()
                release.wait(5.0)

        def gpu_bytes(self):
            return self.n * 50  # five times the 10 B/bus estimate

    srv = serve.Server(main, Overrun, hbm_budget=2000, area_slots=1, bytes_per_bus=10.0)
    srv.start()
    srv.submit(serve.Request(serve.AREA, 0.0, FakeGrid(100), rid=1))
    assert started.wait(5.0)
    # 150 buses: 1,500 B estimated now (fits), 150 x 55 B after the first overrun (does not)
    srv.submit(serve.Request(serve.AREA, 0.0, FakeGrid(150), rid=2))
    srv.submit(serve.Request(serve.AREA, 0.0, FakeGrid(20), rid=3))
    release.set()
    t0 = time.time()
    srv.close(timeout=10.0)
    assert time.time() - t0 < 10.0
    done = {r.rid: r for r in srv.completed}
    assert set(done) == {1, 2, 3}
    assert done[1].error is None and srv.estimate_misses >= 1
    assert done[2].error and "budget" in done[2].error
    assert done[3].error is None  # the area behind it still ran


def test_updates_finish_with_the_estimate_on_the_host():
    """Latency runs to the read-back of the state: every update reads it."""
    calls = []

    class Main(FakeEngine):
        def state(self, out=None):
            calls.append("state")
            return np.zeros(2), np.zeros(2)

    srv = serve.Server(Main("main", []), lambda g: None, hbm_budget=1, area_slots=1)
    reqs = [serve.Request(serve.UPDATE, 0.001 * k, (np.array([k], np.int32), np.array([1.0])), rid=k)
            for k in range(5)]
    done = serve.run(srv, reqs, freeze_gc=False)
    assert all(r.error is None for r in done)
    assert calls.count("state") == 5
