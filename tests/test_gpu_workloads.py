"""The full-size workloads the headline metric is quoted on, on the GPU against
the FP64 oracle (V, theta within 1e-6; J within 1e-8 relative):

- 200,000 buses / ~270,000 branches / 2,000,000 measurements with 1 % gross
  errors (+20 sigma): flat-start Huber M-estimation (gamma = 1.5 sigma), both
  precision forms, against the oracle's own flat-start Huber solve;
- 256 independent areas (500-20,000 buses, log-uniform; redundancy ~3) solved
  as one batched model, from a flat start and from the previous solution
  perturbed by N(0, 0.01), against the oracle on the same concatenated model,
  and sampled areas solved alone;
- the serving loop on the resident 200k-bus model: Poisson updates of 10 % of
  z at 200/s, a breaker operation per 50 updates and new area solves beside
  them; sampled updates are checked against the oracle at the telemetry and
  breaker status they saw, and recorded areas against the oracle too.
Reference semantics: wls_solver.cu:221-431 (loop, modes, Huber), sle_engine.cu:609-682.
"""
import time

import numpy as np
import pytest

from oracle.oracle import Oracle
from state_estimation_with_cuda_amd import gridgen as gg
from state_estimation_with_cuda_amd import serve
from tests.helpers import engine_for

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(900)]

STATE_TOL = 1e-6
J_TOL = 1e-8
GAMMA = 1.5

_CACHE = {}


def grid200k():
    if "g" not in _CACHE:
        _CACHE["g"] = gg.workload_200k()
    return _CACHE["g"]


def oracle_wls(g, key, tol=1e-10):
    """The oracle's flat-start WLS optimum of g (cached under key)."""
    if key not in _CACHE:
        t0 = time.time()
        o = Oracle(g)
        r = o.solve(mode=1, max_iter=60, tol=tol, flat_start=True)
        assert r["status"] == 0, r
        V, th = o.state()
        J = o.solve(mode=1, max_iter=1, tol=1.0)["objective"]
        print(f"[oracle {key}] iters={r['iterations']} {time.time() - t0:.1f}s", flush=True)
        _CACHE[key] = (V, th, J)
    return _CACHE[key]


def errors(V, th, J, ref):
    oV, oth, oJ = ref
    return float(np.abs(V - oV).max()), float(np.abs(th - oth).max()), abs(J - oJ) / abs(oJ)


def supernode_histogram(e):
    sizes = np.diff(e.ordering()["snode_ptr"])
    edges = [1, 2, 3, 5, 9, 17, 33, 65]
    top = max(int(sizes.max()) + 1, edges[-1] + 1)
    h = np.histogram(sizes, bins=edges + [top])[0]
    return {f"{lo}-{hi - 1}": int(c) for lo, hi, c in zip(edges, edges[1:] + [top], h) if c}


@pytest.mark.parametrize("fp64", [True, False])
def test_huber_200k_flat_start_matches_oracle(fp64):
    g = grid200k()
    if "huber" not in _CACHE:
        t0 = time.time()
        o = Oracle(g)
        r = o.solve(mode=1, max_iter=150, tol=1e-9, robust=2, gamma=GAMMA, flat_start=True)
        assert r["status"] == 0, r
        V, th = o.state()
        _CACHE["huber"] = (V, th, r["objective"])
        print(f"[oracle huber200k] iters={r['iterations']} J={r['objective']:.9e} {time.time() - t0:.1f}s",
              flush=True)
    e = engine_for(g, double_precision=fp64)
    t0 = time.time()
    r = e.solve(mode=1, max_iterations=150, tolerance=1e-8, robust=True, gamma=GAMMA, flat_start=True)
    wall = time.time() - t0
    V, th = e.state()
    dV, dth, dJ = errors(V, th, r["objective"], _CACHE["huber"])
    print(f"[huber200k fp64={fp64}] status={r['status']} iters={r['iterations']} refactors={r['refactors']} "
          f"gpu_ms={r['gpu_ms']:.1f} wall={wall:.2f}s dV={dV:.2e} dth={dth:.2e} dJ={dJ:.2e} "
          f"supernodes={supernode_histogram(e)}", flush=True)
    assert r["status"] == 0, r
    assert dV < STATE_TOL and dth < STATE_TOL, (dV, dth)
    assert dJ < J_TOL, dJ
    e.close()


def areas256():
    if "areas" not in _CACHE:
        sizes = gg.area_sizes(256, seed=3)
        areas = [gg.workload_area(int(n), seed=1000 + k) for k, n in enumerate(sizes)]
        _CACHE["areas"] = (areas, gg.concat(areas, name="areas256"))
    return _CACHE["areas"]


@pytest.mark.parametrize("fp64", [True, False])
def test_256_batched_areas_match_oracle(fp64):
    areas, g = areas256()
    ref = oracle_wls(g, "areas256")
    e = engine_for(g, double_precision=fp64)
    r = e.solve(mode=1, max_iterations=50, tolerance=1e-8, flat_start=True)
    V, th = e.state()
    dV, dth, dJ = errors(V, th, r["objective"], ref)
    print(f"[areas256 fp64={fp64} flat] n={g.n} m={g.m} status={r['status']} iters={r['iterations']} "
          f"gpu_ms={r['gpu_ms']:.1f} dV={dV:.2e} dth={dth:.2e} dJ={dJ:.2e}", flush=True)
    assert r["status"] == 0 and dV < STATE_TOL and dth < STATE_TOL and dJ < J_TOL
    # hot start: the previous solution perturbed by N(0, 0.01) per state element
    e.set_state(*gg.hot_start(g, V, th, 0.01, seed=50))
    r = e.solve(mode=1, max_iterations=50, tolerance=1e-8)
    V, th = e.state()
    dV, dth, dJ = errors(V, th, r["objective"], ref)
    print(f"[areas256 fp64={fp64} hot] status={r['status']} iters={r['iterations']} gpu_ms={r['gpu_ms']:.1f} "
          f"dV={dV:.2e} dth={dth:.2e} dJ={dJ:.2e}", flush=True)
    assert r["status"] == 0 and dV < STATE_TOL and dth < STATE_TOL and dJ < J_TOL
    e.close()
    # sampled areas alone: the smallest, a middle one and the largest
    offs = np.cumsum([0] + [a.n for a in areas])
    order = np.argsort([a.n for a in areas])
    for k in (order[0], order[len(order) // 2], order[-1]):
        a, b0 = areas[k], offs[k]
        ea = engine_for(a, double_precision=fp64)
        ra = ea.solve(mode=1, max_iterations=50, tolerance=1e-8, flat_start=True)
        Va, tha = ea.state()
        ea.close()
        assert ra["status"] == 0
        assert np.abs(Va - V[b0:b0 + a.n]).max() < STATE_TOL and np.abs(tha - th[b0:b0 + a.n]).max() < STATE_TOL


@pytest.mark.parametrize("fp64", [True, False])
def test_serving_200k_sampled_requests_match_oracle(fp64):
    g = grid200k()
    rng = np.random.default_rng(4)
    pool = [gg.workload_area(n, seed=7000 + j) for j, n in enumerate((6_000, 15_000, 9_000))]
    main = engine_for(g, double_precision=fp64)
    assert main.solve(mode=1, max_iterations=30, tolerance=1e-8, flat_start=True)["status"] == 0
    wl = serve.make_workload(g, rng, rate=200.0, duration=5.0, toggle_every=50, area_rate=1.0, area_pool=pool,
                             record_every=250, replays=3, record_areas=2)
    srv = serve.Server(main, lambda a: engine_for(a, double_precision=fp64), hbm_budget=4 << 30, area_slots=3,
                       tol=1e-6, max_iterations=10, toggle_mode=0)
    done = serve.run(srv, wl.requests)
    s = serve.summarize(done, warmup=1.0)
    print(f"[serve200k fp64={fp64}] {({k: {q: s[k][q] for q in ('count', 'converged', 'p50_ms', 'p99_ms')} for k in s})}",
          flush=True)
    assert len(done) == len(wl.requests)
    bad = [(r.kind, r.rid, r.status, r.error) for r in done if r.error is not None or r.status != 0]
    assert not bad, bad[:5]
    checked = 0
    for r in done:
        if r.kind == serve.UPDATE and r.rid in wl.replay:
            z, status = wl.replay[r.rid]
            g2 = gg.Grid(**{**g.__dict__})
            g2.z = z
            g2.status = g.status.copy()
            for k, st in status.items():
                g2.status[k] = st
            ref = oracle_wls(g2, f"serve-{r.rid}")
            dV, dth = np.abs(r.result[0] - ref[0]).max(), np.abs(r.result[1] - ref[1]).max()
            print(f"[serve200k fp64={fp64}] update {r.rid} breakers={status} dV={dV:.2e} dth={dth:.2e}", flush=True)
            assert dV < STATE_TOL and dth < STATE_TOL, (r.rid, dV, dth)
            checked += 1
        elif r.kind == serve.AREA and r.record:
            ref = oracle_wls(r.payload, f"area-{r.payload.n}")
            dV, dth = np.abs(r.result[0] - ref[0]).max(), np.abs(r.result[1] - ref[1]).max()
            assert dV < STATE_TOL and dth < STATE_TOL, (r.rid, dV, dth)
            checked += 1
    assert checked >= 4, checked
    main.close()
