"""The allocation-free solve (sle_ext_SolveNoAlloc) against sle_ext_Solve.

Twin engines load the same model and telemetry and take the same solves; one
takes each REALTIME update through sle_ext_Solve, the other through
sle_ext_SolveNoAlloc.  States, objective, status and iteration counts must
agree (the same kernels run, so to rounding), the allocation-free handle's
device memory must not change, and it must never capture a graph.  Calls go
larger model first, then a smaller one (the shared scratch array still holds
the larger call's partials), then the larger one again.  Requests that would
need an allocation are refused with SLE_ERROR_INVALID_ARGUMENT.
"""
import numpy as np
import pytest

from state_estimation_with_cuda_amd import gridgen as gg
from tests.helpers import engine_for

pytestmark = pytest.mark.gpu

INVALID_ARGUMENT = -4


def _pair(g, fp64):
    ref = engine_for(g, double_precision=fp64)
    na = engine_for(g, double_precision=fp64)
    for e in (ref, na):
        r = e.solve(mode=1, max_iterations=50, tolerance=1e-8, flat_start=True)
        assert r["status"] == 0, r
    return ref, na


def _update(ref, na, g, seed, label):
    rng = np.random.default_rng(seed)
    idx = np.sort(rng.choice(g.m, g.m // 10, replace=False)).astype(np.int32)
    val = g.z[idx] + rng.standard_normal(idx.size) * g.sigma[idx]
    before = na.gpu_bytes()
    ref.set_telemetry_indexed(idx, val)
    na.set_telemetry_indexed(idx, val)
    rr = ref.solve(mode=0, max_iterations=20, tolerance=1e-8)
    rn = na.solve(mode=0, max_iterations=20, tolerance=1e-8, no_alloc=True)
    print(f"[{label}] ref: status={rr['status']} iters={rr['iterations']} J={rr['objective']:.9e}; "
          f"no-alloc: rc={rn['rc']} status={rn['status']} iters={rn['iterations']} J={rn['objective']:.9e}",
          flush=True)
    assert rn["rc"] == 0 and rn["status"] == 0, rn
    assert (rn["status"], rn["iterations"], rn["refactors"]) == (rr["status"], rr["iterations"], rr["refactors"])
    assert rn["graph_captures"] == 0 and rn["seq_fallbacks"] == 0, rn
    (Vr, tr), (Vn, tn) = ref.state(), na.state()
    dV, dth = np.abs(Vr - Vn).max(), np.abs(tr - tn).max()
    assert dV < 1e-9 and dth < 1e-9, (dV, dth)
    assert abs(rn["objective"] - rr["objective"]) <= 1e-12 * rr["objective"]
    assert na.gpu_bytes() == before  # nothing allocated or released


@pytest.mark.parametrize("fp64", [True, False])
def test_no_alloc_solve_matches_ext_solve(fp64):
    big = gg.workload_10k()
    small = gg.geometric(2000, seed=11, flow_meas=6000)
    rb, nb_ = _pair(big, fp64)
    rs, ns_ = _pair(small, fp64)
    _update(rb, nb_, big, 1, f"10k fp64={fp64}")
    _update(rs, ns_, small, 2, f"2k fp64={fp64} (smaller, after the larger call)")
    _update(rb, nb_, big, 3, f"10k fp64={fp64} (after the smaller call)")
    # the allocation-free handle then takes a regular update: its replays still
    # see every status (the iteration sequence was carried over)
    rng = np.random.default_rng(4)
    idx = np.sort(rng.choice(big.m, big.m // 10, replace=False)).astype(np.int32)
    nb_.set_telemetry_indexed(idx, big.z[idx])
    r = nb_.solve(mode=0, max_iterations=20, tolerance=1e-8)
    assert r["status"] == 0 and r["seq_fallbacks"] == 0, r
    for e in (rb, nb_, rs, ns_):
        e.close()


def test_no_alloc_refuses_requests_that_need_allocations():
    g = gg.geometric(500, seed=3, flow_meas="all")
    e = engine_for(g, double_precision=True)
    # no cached factor yet
    r = e.solve(mode=0, max_iterations=10, tolerance=1e-8, no_alloc=True)
    assert r["rc"] == INVALID_ARGUMENT and "factor" in e.last_error(), (r, e.last_error())
    assert e.solve(mode=1, max_iterations=50, tolerance=1e-8, flat_start=True)["status"] == 0
    # PRECISION refactors; Huber refactors
    r = e.solve(mode=1, max_iterations=10, tolerance=1e-8, no_alloc=True)
    assert r["rc"] == INVALID_ARGUMENT and "REALTIME" in e.last_error(), e.last_error()
    r = e.solve(mode=0, max_iterations=10, tolerance=1e-8, robust=True, no_alloc=True)
    assert r["rc"] == INVALID_ARGUMENT, r
    # a pending breaker operation needs the topology refresh first (a branch
    # on a cycle, so the grid stays one island)
    import scipy.sparse as sp
    import scipy.sparse.csgraph as cg

    k = 0
    for k in range(g.nb):
        keep = np.arange(g.nb) != k
        A = sp.coo_matrix((np.ones(keep.sum()), (g.br_from[keep], g.br_to[keep])), shape=(g.n, g.n))
        if cg.connected_components(A, directed=False)[0] == 1:
            break
    e.set_branch_status(k, 0)
    r = e.solve(mode=0, max_iterations=10, tolerance=1e-8, no_alloc=True)
    assert r["rc"] == INVALID_ARGUMENT and "pending" in e.last_error(), e.last_error()
    # after a regular solve the allocation-free one runs again
    assert e.solve(mode=0, max_iterations=20, tolerance=1e-8)["status"] == 0
    r = e.solve(mode=0, max_iterations=20, tolerance=1e-8, no_alloc=True)
    assert r["rc"] == 0 and r["status"] == 0, r
    e.close()
