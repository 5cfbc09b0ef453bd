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

    def state(self):
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
