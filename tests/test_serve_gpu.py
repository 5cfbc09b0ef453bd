"""Serving loop on the GPU: a resident model takes updates and breaker
operations while new areas are solved beside it under an HBM budget.  Each
result must be the same alone and in the mix (within the 1e-6 tolerance;
area solves are deterministic), and the budget must hold."""
import numpy as np
import pytest

from state_estimation_with_cuda_amd import gridgen as gg
from state_estimation_with_cuda_amd import serve, sle

pytestmark = pytest.mark.gpu


def engine(grid):
    e = sle.Engine(double=True)
    e.load_grid(grid)
    e.set_config(use_double_precision=1)
    e.upload()
    e.set_telemetry_f64(grid.z)
    return e


def test_serving_mix_matches_requests_alone():
    g = gg.geometric(3000, seed=61, flow_meas="all")
    main = engine(g)
    assert main.solve(mode=1, max_iterations=50, tolerance=1e-8, flat_start=True)["status"] == 0
    rng = np.random.default_rng(5)
    areas = [gg.workload_area(n, seed=80 + j) for j, n in enumerate((900, 1500, 700, 1200))]
    cal = engine(areas[0])
    bpb = 1.25 * cal.gpu_bytes() / areas[0].n
    cal.close()
    budget = int(bpb * 2300)  # about two areas at a time
    k = g.m // 10
    z = g.z.copy()
    reqs, snaps = [], {}
    t = serve.poisson_arrivals(300.0, 0.25, rng)
    for q, a in enumerate(t):
        idx = np.sort(rng.choice(g.m, k, replace=False)).astype(np.int32)
        val = g.z[idx] + rng.normal(0, 1, k) * g.sigma[idx]
        z[idx] = val
        rec = q % 10 == 5
        reqs.append(serve.Request(serve.UPDATE, float(a), (idx, val), rid=q, record=rec))
        if rec:
            snaps[q] = z.copy()
    for j, grid in enumerate(areas):
        reqs.append(serve.Request(serve.AREA, 0.02 * j, grid, rid=1000 + j, record=True))
    print(f"[serve] updates={len(t)} areas={len(areas)} budget={budget}", flush=True)
    srv = serve.Server(main, engine, hbm_budget=budget, area_slots=3, bytes_per_bus=bpb, tol=1e-7, max_iterations=20)
    done = serve.run(srv, reqs)
    assert len(done) == len(reqs)
    assert all(r.error is None and r.status == 0 for r in done), [(r.kind, r.rid, r.error, r.status) for r in done]
    assert srv.peak_used <= budget
    for r in done:
        if r.kind == serve.AREA:
            e = engine(r.payload)
            e.solve(mode=1, max_iterations=50, tolerance=1e-7, flat_start=True)
            V, th = e.state()
            e.close()
            assert np.abs(V - r.result[0]).max() < 1e-12 and np.abs(th - r.result[1]).max() < 1e-12
        elif r.record:
            main.set_telemetry_f64(snaps[r.rid])
            main.set_state(*r.before[:2])
            main.solve(mode=0, max_iterations=20, tolerance=1e-7)
            V, th = main.state()
            assert np.abs(V - r.result[0]).max() < 1e-6 and np.abs(th - r.result[1]).max() < 1e-6
    s = serve.summarize(done)
    print(f"[serve] {s}", flush=True)
    main.close()
