"""Serving loop over the sle_* C-ABI: requests admitted as they arrive.

One resident model (its gain factor and hot-start state stay in HBM) takes
incremental telemetry updates and breaker operations in arrival order; full
solves of new areas run beside it in a pool of slots, each on its own engine
(its own HIP stream), admitted first-come first-served while their estimated
HBM footprint fits the budget.  Work of the resident model and of the areas
overlaps on the device.  Every request records arrival, start and completion
times; latency is completion (the estimate read back to the host) minus
arrival.

The reference engine has no serving layer (its ABI is synchronous and one
handle is not to be used from two threads, sle_api.h:60-63); this module keeps
that contract -- each handle is driven by exactly one thread -- and adds the
admission and overlap around it.
"""
from __future__ import annotations

import collections
import dataclasses
import gc
import threading
import time
from typing import Any, Callable, Optional

import numpy as np

UPDATE, TOGGLE, AREA = "update", "toggle", "area"


@dataclasses.dataclass
class Request:
    kind: str
    arrival: float  # seconds after the start of the run
    payload: Any = None  # update: (idx, values); toggle: (branch, status[, idx, values]); area: grid
    rid: int = 0
    record: bool = False  # keep the result (and, for updates, the starting point) for replay
    submitted: Optional[float] = None
    start: Optional[float] = None
    done: Optional[float] = None
    status: Optional[int] = None
    iterations: int = 0
    error: Optional[str] = None
    result: Any = None
    before: Any = None  # updates with record: (V, theta, breaker status) before the request

    @property
    def latency(self) -> float:
        return (self.done - self.arrival) if self.done is not None else float("nan")


class Server:
    """Continuous admission of requests.

    main: engine of the resident model (sle.Engine-like: set_telemetry_indexed,
          set_branch_status, solve, state).
    make_area_engine(grid) -> engine with the area loaded, uploaded and its
          telemetry set; the server closes it after the solve.
    hbm_budget: bytes available to area slots; an area is admitted when its
          estimate fits beside the areas already running.
    """

    def __init__(self, main, make_area_engine: Callable, hbm_budget: int, area_slots: int = 4,
                 bytes_per_bus: float = 16384.0, tol: float = 1e-6, max_iterations: int = 10,
                 clock: Callable[[], float] = time.perf_counter, toggle_mode: int = 0):
        self.main = main
        self.make_area_engine = make_area_engine
        self.budget = int(hbm_budget)
        self.slots = int(area_slots)
        self.bytes_per_bus = float(bytes_per_bus)
        self.tol = tol
        self.max_iterations = max_iterations
        self.clock = clock
        # solve mode after a breaker operation: 1 PRECISION (refactor every
        # step), 0 REALTIME (the invalidated factor is rebuilt once, then reused
        # and refreshed only if it contracts poorly)
        self.toggle_mode = toggle_mode
        self._cv = threading.Condition()
        self._main_q: collections.deque = collections.deque()
        self._area_q: collections.deque = collections.deque()
        self._used = 0
        self.peak_used = 0
        self.estimate_misses = 0
        self._running_areas = 0
        self._closing = False
        self._t0 = 0.0
        self.completed: list = []
        self._breakers: dict = {}
        self._state_buf = None  # the resident model's estimate, read back into the same arrays
        self._threads = [threading.Thread(target=self._main_loop, name="sle-main", daemon=True)]
        self._threads += [threading.Thread(target=self._area_loop, name=f"sle-area{k}", daemon=True)
                          for k in range(self.slots)]

    # -- control -----------------------------------------------------------
    def start(self):
        self._t0 = self.clock()
        for t in self._threads:
            t.start()
        return self._t0

    def now(self) -> float:
        return self.clock() - self._t0

    def estimate_bytes(self, grid) -> int:
        return int(self.bytes_per_bus * grid.n)

    def submit(self, req: Request):
        req.submitted = self.now()
        with self._cv:
            if req.kind == AREA:
                if self.estimate_bytes(req.payload) > self.budget:
                    req.error = "area larger than the HBM budget"
                    req.done = self.now()
                    self.completed.append(req)
                    return
                self._area_q.append(req)
            else:
                self._main_q.append(req)
            self._cv.notify_all()

    def close(self, timeout: float = 600.0):
        """Wait until every submitted request is done, then stop the workers."""
        with self._cv:
            self._closing = True
            self._cv.notify_all()
        for t in self._threads:
            t.join(timeout)

    # -- workers -----------------------------------------------------------
    def _main_loop(self):
        while True:
            with self._cv:
                while not self._main_q and not self._closing:
                    self._cv.wait()
                if not self._main_q:
                    return
                req = self._main_q.popleft()
            req.start = self.now()
            try:
                if req.kind == UPDATE:
                    idx, val = req.payload
                    if req.record:
                        V, th = self.main.state()
                        req.before = (V, th, dict(self._breakers))
                    self.main.set_telemetry_indexed(idx, val)
                else:  # (branch, status[, telemetry indices, values that change with it])
                    k, st = req.payload[:2]
                    if len(req.payload) > 2 and len(req.payload[2]):
                        self.main.set_telemetry_indexed(req.payload[2], req.payload[3])
                    self.main.set_branch_status(k, st)
                    self._breakers[k] = st
                # updates reuse the cached factor (REALTIME); a breaker operation
                # moves the optimum and invalidates the factor (toggle_mode)
                mode = 0 if req.kind == UPDATE else self.toggle_mode
                r = self.main.solve(mode=mode, max_iterations=self.max_iterations, tolerance=self.tol)
                req.status, req.iterations = r["status"], r["iterations"]
                # the request is done when its estimate is on the host
                if req.record:
                    req.result = self.main.state()
                else:
                    self._state_buf = self.main.state(out=self._state_buf)
            except Exception as exc:  # the request fails, the loop goes on
                req.error = f"{type(exc).__name__}: {exc}"
            req.done = self.now()
            with self._cv:
                self.completed.append(req)

    def _area_loop(self):
        while True:
            with self._cv:
                while True:
                    if self._area_q:
                        need = self.estimate_bytes(self._area_q[0].payload)
                        if need > self.budget:
                            # the estimate grew past the whole budget after this
                            # request was queued (an overrun raised bytes_per_bus):
                            # it can never be admitted, so it fails instead of
                            # holding up every request behind it
                            head = self._area_q.popleft()
                            head.error = "area larger than the HBM budget"
                            head.done = self.now()
                            self.completed.append(head)
                            self._cv.notify_all()
                            continue
                        if self._used + need <= self.budget:
                            break
                    elif self._closing:
                        return
                    self._cv.wait()
                req = self._area_q.popleft()  # first come, first served
                need = self.estimate_bytes(req.payload)
                self._used += need
                self.peak_used = max(self.peak_used, self._used)
                self._running_areas += 1
            req.start = self.now()
            e = None
            try:
                e = self.make_area_engine(req.payload)
                r = e.solve(mode=1, max_iterations=50, tolerance=self.tol, flat_start=True)
                req.status, req.iterations = r["status"], r["iterations"]
                state = e.state()  # the request is done when its estimate is on the host
                req.result = state if req.record else None
                used = e.gpu_bytes() if hasattr(e, "gpu_bytes") else 0
                if used > need:  # learn from the overrun: later admissions use the larger figure
                    with self._cv:
                        self.estimate_misses += 1
                        self.bytes_per_bus = max(self.bytes_per_bus, 1.1 * used / max(req.payload.n, 1))
            except Exception as exc:
                req.error = f"{type(exc).__name__}: {exc}"
            finally:
                if e is not None:
                    e.close()
            req.done = self.now()
            with self._cv:
                self._used -= need
                self._running_areas -= 1
                self.completed.append(req)
                self._cv.notify_all()


def non_bridge_branches(g, count: int, rng: np.random.Generator) -> list:
    """Branches whose opening keeps the grid connected (a breaker operation on
    one of them changes no island)."""
    import scipy.sparse as sp
    import scipy.sparse.csgraph as cg

    out = []
    for k in rng.permutation(g.nb):
        keep = np.ones(g.nb, bool)
        keep[k] = False
        A = sp.coo_matrix((np.ones(keep.sum()), (g.br_from[keep], g.br_to[keep])), shape=(g.n, g.n))
        if cg.connected_components(A, directed=False)[0] == 1:
            out.append(int(k))
            if len(out) == count:
                return out
    return out


@dataclasses.dataclass
class Workload:
    """A request stream on a resident model, with what is needed to replay or
    check any recorded request: the telemetry and breaker status it saw."""
    requests: list
    replay: dict  # update rid -> (z after the request, {branch: status} at that request)
    breakers: dict  # every branch operated, with its last status


def make_workload(g, rng: np.random.Generator, rate: float, duration: float, frac: float = 0.10,
                  toggle_every: int = 50, area_rate: float = 0.0, area_pool=(), record_every: int = 500,
                  replays: int = 4, record_areas: int = 3) -> Workload:
    """Poisson arrivals of incremental updates (each re-measures a random `frac`
    of z: the base value h(x_true) + noise plus fresh N(0, sigma^2)), a breaker
    operation after every `toggle_every` updates (alternately opening and
    closing a non-bridge branch; the injection readings at its ends follow the
    new topology at the true state and travel with the operation), and full
    solves of new areas drawn from `area_pool` at `area_rate` per second.
    Every `record_every`-th update (offset by half) keeps its result, and the
    first `replays` of them keep their telemetry and breaker status."""
    from . import gridgen

    t_upd = poisson_arrivals(rate, duration, rng)
    t_area = poisson_arrivals(area_rate, duration, rng) if area_rate > 0 and len(area_pool) else np.zeros(0)
    brks = non_bridge_branches(g, max(1, len(t_upd) // 100 + 1), rng) if toggle_every else []
    k = max(1, int(round(frac * g.m)))
    z = g.z.copy()  # the telemetry as the resident model sees it after each request
    zb = g.z.copy()  # the telemetry base for the current breaker status: h(x_true) + the original noise
    noise = g.z - g.z_true
    inj = np.isin(g.m_type, [gridgen.P_INJ, gridgen.Q_INJ, gridgen.P_PSEUDO, gridgen.Q_PSEUDO])
    status_arr = g.status.copy()
    status: dict = {}
    reqs, replay = [], {}
    open_brk = None
    for q, t in enumerate(t_upd):
        idx = np.sort(rng.choice(g.m, k, replace=False)).astype(np.int32)
        val = zb[idx] + rng.normal(0.0, 1.0, k) * g.sigma[idx]
        z[idx] = val
        rec = q % record_every == record_every // 2
        reqs.append(Request(UPDATE, float(t), (idx, val), rid=q, record=rec))
        if rec and len(replay) < replays:
            replay[q] = (z.copy(), dict(status))
        if toggle_every and (q + 1) % toggle_every == 0 and brks:
            if open_brk is None:
                open_brk = brks[(q // 100) % len(brks)]
                op = (open_brk, 0)
            else:
                op = (open_brk, 1)
                open_brk = None
            status[op[0]] = op[1]
            status_arr[op[0]] = op[1]
            gt = dataclasses.replace(g, status=status_arr)
            ends = (g.br_from[op[0]], g.br_to[op[0]])
            rows = np.nonzero(inj & np.isin(g.m_loc, ends))[0].astype(np.int32)
            zb[rows] = gridgen.measurement_function(gt, g.V_true, g.th_true, m_type=g.m_type[rows],
                                                    m_loc=g.m_loc[rows], m_end=g.m_end[rows]) + noise[rows]
            z[rows] = zb[rows]
            reqs.append(Request(TOGGLE, float(t) + 1e-6, op + (rows, zb[rows].copy()), rid=100_000 + q))
    for j, t in enumerate(t_area):
        reqs.append(Request(AREA, float(t), area_pool[j % len(area_pool)], rid=200_000 + j, record=j < record_areas))
    return Workload(reqs, replay, status)


def poisson_arrivals(rate: float, duration: float, rng: np.random.Generator) -> np.ndarray:
    t, out = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration:
            return np.array(out)
        out.append(t)


def run(server: Server, requests: list, sleep: Callable[[float], None] = time.sleep, freeze_gc: bool = True) -> list:
    """Release each request at its arrival time, then wait for all of them.

    freeze_gc: move every object alive at the start out of the cyclic garbage
    collector's reach for the run (a full collection over a large model's
    Python objects stops all threads for a noticeable time)."""
    requests = sorted(requests, key=lambda r: r.arrival)
    if freeze_gc:
        gc.collect()
        gc.freeze()
    try:
        server.start()
        for req in requests:
            wait = req.arrival - server.now()
            if wait > 0:
                sleep(wait)
            server.submit(req)
        server.close()
    finally:
        if freeze_gc:
            gc.unfreeze()
    return server.completed


def summarize(done: list, warmup: float = 0.0) -> dict:
    out = {}
    for kind in (UPDATE, TOGGLE, AREA):
        rs = [r for r in done if r.kind == kind and r.arrival >= warmup and r.done is not None]
        lat = np.array([r.latency for r in rs if r.error is None]) * 1e3
        its = np.array([r.iterations for r in rs if r.error is None])
        out[kind] = {
            "count": len(rs),
            "failed": sum(1 for r in rs if r.error is not None),
            "converged": sum(1 for r in rs if r.status == 0),
            "status_counts": {str(k): int(v) for k, v in collections.Counter(r.status for r in rs).items()},
            "mean_iterations": float(its.mean()) if its.size else None,
            "max_iterations": int(its.max()) if its.size else None,
            "p50_ms": float(np.percentile(lat, 50)) if lat.size else None,
            "p99_ms": float(np.percentile(lat, 99)) if lat.size else None,
            "max_ms": float(lat.max()) if lat.size else None,
        }
    return out
