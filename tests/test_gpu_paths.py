"""GPU paths beyond the plain solve, through the sle_* C-ABI against the oracle.

- measurement types past V/P/Q: branch current magnitude with its exact
  derivative, PMU angles and a phase shifter (MATPOWER convention);
- failure semantics of the Gauss-Newton loop (wls_solver.cu:221-372,
  sle_api.cu:971-984): NaN telemetry diverges, with two flat-start retries in
  PRECISION and none in REALTIME; an unobservable gain is SINGULAR; the REALTIME
  time limit stops the loop at MAX_ITERATIONS;
- the zero-impedance clamp r^2 + x^2 >= 1e-12 (network_model.cu:602-603);
- a breaker that splits an island: the cut-off part gets its own angle
  reference (symbolic redo on the device), and both states match the oracle;
- observability (sle_CheckObservability: islands, then the gain's pivots);
- meters (network_model.cu:312-455): channels map to measurements, and the
  estimate and residual of a channel are those of its measurement;
- captured iteration graphs survive other engines being created and closed.
"""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.csgraph as cg

from oracle.oracle import Oracle
from state_estimation_with_cuda_amd import gridgen as gg
from state_estimation_with_cuda_amd import sle
from tests.helpers import GRIDS, engine_for, rel_err
from tests.test_oracle_pins import variant14

pytestmark = pytest.mark.gpu

STATE_TOL = 1e-6
J_TOL = 1e-8

# sle_api.h status codes and convergence states
OK, NOT_OBSERVABLE, SINGULAR_MATRIX, DIVERGED, MAX_ITERATIONS = 0, -6, -7, -8, -9
CONVERGED, ST_MAX_ITER, ST_DIVERGED, ST_SINGULAR = 0, 1, 2, 3


def oracle_optimum(g, tol=1e-11):
    o = Oracle(g)
    r = o.solve(mode=1, max_iter=100, tol=tol, flat_start=True)
    assert r["status"] == 0, r
    V, th = o.state()
    J = o.solve(mode=1, max_iter=1, tol=1.0)["objective"]
    return V, th, J


def assert_matches(e, r, g, label):
    oV, oth, oJ = oracle_optimum(g)
    V, th = e.state()
    dV, dth = np.abs(V - oV).max(), np.abs(th - oth).max()
    dJ = abs(r["objective"] - oJ) / oJ
    print(f"[{label}] iters={r['iterations']} dV={dV:.2e} dth={dth:.2e} dJ={dJ:.2e}", flush=True)
    assert dV < STATE_TOL and dth < STATE_TOL, (dV, dth)
    assert dJ < J_TOL, dJ


@pytest.mark.parametrize("fp64", [True, False])
def test_current_angle_and_phase_shifter_rows_match_oracle(fp64):
    """I_MAG rows (exact derivative of |S|/V), V_ANGLE rows and a phase shifter:
    h(x) and H values at a perturbed state, then the estimate."""
    g = variant14()
    assert (g.m_type == gg.I_MAG).any() and (g.m_type == gg.V_ANGLE).any() and (g.shift != 0).any()
    o = Oracle(g)
    e = engine_for(g, double_precision=fp64)
    rng = np.random.default_rng(8)
    V = g.V_true + rng.normal(0, 0.02, g.n)
    th = g.th_true + rng.normal(0, 0.02, g.n)
    th[g.slack] = 0
    e.set_state(V, th)
    ev, oe = e.evaluate(), o.eval(V, th)
    assert rel_err(ev["h"], oe["h"]) < 1e-12
    hv, _ = e.jacobian_gain()
    assert rel_err(hv, o.jacobian(V, th)) < 1e-12  # Jacobian values are FP64 in both forms
    r = e.solve(mode=1, max_iterations=50, tolerance=1e-8, flat_start=True)
    assert r["status"] == CONVERGED, r
    assert_matches(e, r, g, f"variant14 fp64={fp64}")
    e.close()


@pytest.mark.parametrize("mode", [1, 0])
def test_nan_telemetry_diverges_with_retries_in_precision_only(mode):
    g = GRIDS["grid14"]()
    e = engine_for(g, double_precision=True)
    assert e.solve(mode=1, max_iterations=50, tolerance=1e-8, flat_start=True)["status"] == CONVERGED
    z = g.z.copy()
    z[5] = np.nan
    e.set_telemetry_f64(z)
    r = e.solve(mode=mode, max_iterations=20, tolerance=1e-8)
    print(f"[nan mode={mode}] {r}", flush=True)
    assert r["status"] == ST_DIVERGED and r["rc"] == DIVERGED, r
    assert r["retries"] == (2 if mode == 1 else 0), r
    # good telemetry again: a flat start recovers the oracle's optimum
    e.set_telemetry_f64(g.z)
    r = e.solve(mode=1, max_iterations=50, tolerance=1e-8, flat_start=True)
    assert r["status"] == CONVERGED and r["rc"] == OK, r
    assert_matches(e, r, g, "after nan")
    e.close()


def unobservable14():
    """14-bus grid with |V| at every bus but P/Q injections at three buses only
    and no flows: 20 measurements for 27 states, so some angles cannot be
    resolved and the gain is singular."""
    g = gg.grid14(0)
    keep = (g.m_type == gg.V_MAG) | (np.isin(g.m_type, [gg.P_INJ, gg.Q_INJ]) & (g.m_loc < 3))
    for f in ("m_type", "m_loc", "m_end", "sigma", "z", "z_true"):
        setattr(g, f, getattr(g, f)[keep])
    return g


def test_unobservable_gain_is_singular_and_not_observable():
    g = unobservable14()
    assert g.m < g.ns  # fewer measurements than states
    e = engine_for(g, double_precision=True)
    obs, isl = C.c_int32(-1), C.c_int32(-1)
    assert e.L.sle_CheckObservability(e.h, C.byref(obs), C.byref(isl)) == OK
    assert isl.value == 1 and obs.value == 0
    r = e.solve(mode=1, max_iterations=20, tolerance=1e-8, flat_start=True)
    print(f"[unobservable] {r}", flush=True)
    assert r["status"] == ST_SINGULAR and r["rc"] == SINGULAR_MATRIX, r
    e.close()


def test_observable_grid_reports_observable():
    g = GRIDS["geo2k"]()
    e = engine_for(g, double_precision=True)
    obs, isl = C.c_int32(-1), C.c_int32(-1)
    assert e.L.sle_CheckObservability(e.h, C.byref(obs), C.byref(isl)) == OK
    assert isl.value == 1 and obs.value == 1
    # the check leaves the engine able to solve
    r = e.solve(mode=1, max_iterations=50, tolerance=1e-8, flat_start=True)
    assert r["status"] == CONVERGED
    e.close()


def test_realtime_time_limit_stops_at_max_iterations():
    g = GRIDS["geo10k"]()
    e = engine_for(g, double_precision=True)
    assert e.solve(mode=1, max_iterations=50, tolerance=1e-8, flat_start=True)["status"] == CONVERGED
    rng = np.random.default_rng(3)
    idx = np.arange(0, g.m, 3, dtype=np.int32)
    e.set_telemetry_indexed(idx, g.z[idx] + 5.0 * g.sigma[idx] * rng.standard_normal(idx.size))
    # a limit far below one iteration's time: the loop stops after its first
    # iteration (wls_solver.cu:313-323) although the tolerance is not reached
    r = e.solve(mode=0, max_iterations=50, tolerance=1e-14, time_limit_ms=1e-3)
    print(f"[time limit] {r}", flush=True)
    assert r["status"] == ST_MAX_ITER and r["rc"] == MAX_ITERATIONS, r
    assert r["iterations"] == 1, r
    # without the limit the same request converges to the oracle's optimum
    r = e.solve(mode=0, max_iterations=50, tolerance=1e-8)
    assert r["status"] == CONVERGED, r
    e.close()


@pytest.mark.parametrize("fp64", [True, False])
def test_zero_impedance_branch_clamp(fp64):
    """r = x = 0 clamps r^2 + x^2 to 1e-12 (g = b = 0: no admittance); a branch
    with r = 0, x = 1e-7 clamps to b = -1e5 pu.  Ybus values, and for the
    second case the estimate, match the oracle."""
    g0 = gg.geometric(300, seed=17, flow_meas="all")
    for k_r, k_x in ((0.0, 0.0), (0.0, 1e-7)):
        g = gg.Grid(**{**g0.__dict__})
        g.r, g.x, g.b = g0.r.copy(), g0.x.copy(), g0.b.copy()
        k = 7
        g.r[k], g.x[k], g.b[k] = k_r, k_x, 0.0
        g.z_true = gg.measurement_function(g, g.V_true, g.th_true)
        g.z = g.z_true + np.random.default_rng(2).normal(0, 1, g.m) * g.sigma
        o = Oracle(g)
        e = engine_for(g, double_precision=fp64)
        _, _, G, B = e.ybus()
        _, _, oG, oB = o.ybus()
        assert rel_err(G, oG) < 1e-12 and rel_err(B, oB) < 1e-12
        if k_x > 0:
            assert np.abs(B).max() > 1e4  # the clamped series susceptance is in Ybus
            r = e.solve(mode=1, max_iterations=50, tolerance=1e-9, flat_start=True)
            assert r["status"] == CONVERGED, r
            assert_matches(e, r, g, f"zero impedance fp64={fp64}")
        e.close()


def _bridge_and_side(g):
    """A bridge of the branch graph and the buses its removal cuts off from the slack."""
    A = sp.coo_matrix((np.ones(g.nb), (g.br_from, g.br_to)), shape=(g.n, g.n))
    best = None
    for k in range(g.nb):
        keep = np.ones(g.nb, bool)
        keep[k] = False
        Ak = sp.coo_matrix((np.ones(keep.sum()), (g.br_from[keep], g.br_to[keep])), shape=(g.n, g.n))
        nc, lab = cg.connected_components(Ak, directed=False)
        if nc == 2:
            side = np.nonzero(lab != lab[g.slack])[0]
            if 5 <= side.size and (best is None or side.size > best[1].size):
                best = (k, side)
    assert best is not None and A.nnz
    return best


@pytest.mark.parametrize("fp64", [True, False])
def test_bridge_opening_splits_island_on_device(fp64):
    """Open a bridge: the cut-off part (with a SLACK-typed bus of its own, which
    is an ordinary state while the grid is one island) becomes an island with its
    own angle reference.  Symbolic results and states match the oracle before,
    after the split and after closing the bridge again."""
    g = gg.geometric(1200, seed=5, flow_meas="all")
    k, side = _bridge_and_side(g)
    g.bus_type = g.bus_type.copy()
    g.bus_type[int(side.max())] = gg.BUS_SLACK
    e = engine_for(g, double_precision=fp64)
    o = Oracle(g)
    for step, st in enumerate((1, 0, 1)):
        if step:
            e.set_branch_status(k, st)
            o.set_branch_status(k, st)
        r = e.solve(mode=1, max_iterations=50, tolerance=1e-9, flat_start=True)
        assert r["status"] == CONVERGED, (step, r)
        assert e.sizes()["ns"] == o.ns
        es, os_ = e.ordering(), o.symbolic()
        for key in ("perm", "parent", "snode_ptr"):
            assert np.array_equal(es[key], os_[key]), (step, key)
        g2 = gg.Grid(**{**g.__dict__})
        g2.status = g.status.copy()
        g2.status[k] = st
        assert_matches(e, r, g2, f"bridge step {step} status {st} fp64={fp64} ns={o.ns}")
    e.close()


def test_meter_channels_estimate_and_residual():
    """A multimeter on a branch end adds V, P, Q and I channels as measurements
    (ids <meter>.V/.kW/.kVAR/.A); readings set through the meter API feed the
    solve, and the meter's estimate and residual are those of the oracle at its
    optimum for the same measurement set."""
    g = gg.grid14(0)
    k = 3
    bus = int(g.br_from[k])
    extra_t = np.array([gg.V_MAG, gg.P_FLOW, gg.Q_FLOW, gg.I_MAG], np.int32)
    extra_l = np.array([bus, k, k, k], np.int32)
    sig = np.array([0.004, 0.01, 0.01, 0.01])
    g2 = gg.Grid(**{**g.__dict__})
    g2.m_type = np.concatenate([g.m_type, extra_t])
    g2.m_loc = np.concatenate([g.m_loc, extra_l])
    g2.m_end = np.concatenate([g.m_end, np.zeros(4, np.int32)])
    g2.sigma = np.concatenate([g.sigma, sig])
    g2.z_true = gg.measurement_function(g2, g.V_true, g.th_true)
    g2.z = np.concatenate([g.z, g2.z_true[-4:] + sig * np.array([0.5, -1.0, 0.7, 0.3])])

    e = sle.Engine(double=True)
    e.load_grid(g)
    L, S, h = e.L, e.S, e.h
    info = S["MeterInfo"](b"MT1", 1, f"B{bus}".encode(), f"L{k}".encode(), 0, 1.0, 1.0, 0.004, 0.01, 0.01)
    assert L.sle_AddMeter(h, C.byref(info)) == OK
    assert L.sle_GetMeterCount(h) == 1
    e.m = g.m + 4
    e.upload()
    e.set_telemetry_f64(np.concatenate([g.z, np.zeros(4)]))
    channels = [b"V", b"kW", b"kVAR", b"A"]
    for ch, val in zip(channels, g2.z[-4:]):
        assert L.sle_UpdateMeterReading(h, b"MT1", ch, val) == OK
    assert L.sle_UpdateMeterReading(h, b"MT1", b"bogus", 1.0) != OK
    r = e.solve(mode=1, max_iterations=50, tolerance=1e-9, flat_start=True)
    assert r["status"] == CONVERGED, r
    assert_matches(e, r, g2, "meter")
    o = Oracle(g2)
    o.solve(mode=1, max_iter=100, tol=1e-11, flat_start=True)
    oh = o.eval(*o.state())["h"][-4:]
    for j, ch in enumerate(channels):
        rd, est, res = C.c_double(), C.c_double(), C.c_double()
        assert L.sle_GetMeterReading(h, b"MT1", ch, C.byref(rd)) == OK
        assert L.sle_GetMeterEstimate(h, b"MT1", ch, C.byref(est)) == OK
        assert L.sle_GetMeterResidual(h, b"MT1", ch, C.byref(res)) == OK
        assert rd.value == g2.z[g.m + j]
        assert abs(est.value - oh[j]) < 1e-6, (ch, est.value, oh[j])
        assert abs(res.value - (rd.value - est.value)) < 1e-12
    e.close()


@pytest.mark.parametrize("fp64", [True, False])
def test_graphs_survive_other_engines(fp64):
    """An engine's captured iteration graph is kept while other engines are
    created, solved and closed beside it (their allocations are not its own),
    and every replayed iteration's status is seen without the fallback wait."""
    g = GRIDS["geo2k"]()
    e = engine_for(g, double_precision=fp64, use_cuda_graphs=1)
    assert e.solve(mode=1, max_iterations=50, tolerance=1e-8, flat_start=True)["status"] == CONVERGED
    rng = np.random.default_rng(4)
    idx = np.sort(rng.choice(g.m, g.m // 10, replace=False)).astype(np.int32)
    e.set_telemetry_indexed(idx, g.z[idx] + 0.3 * g.sigma[idx])
    r = e.solve(mode=0, max_iterations=20, tolerance=1e-8)
    assert r["status"] == CONVERGED and r["graph_captures"] >= 1 and r["seq_fallbacks"] == 0, r
    for j in range(2):
        a = gg.geometric(500, seed=60 + j, flow_meas="all")
        ea = engine_for(a, double_precision=fp64)
        assert ea.solve(mode=1, max_iterations=50, tolerance=1e-8, flat_start=True)["status"] == CONVERGED
        ea.close()
        e.set_telemetry_indexed(idx, g.z[idx] - 0.2 * (j + 1) * g.sigma[idx])
        r = e.solve(mode=0, max_iterations=20, tolerance=1e-8)
        print(f"[graphs j={j} fp64={fp64}] {r}", flush=True)
        assert r["status"] == CONVERGED and r["refactors"] == 0, r
        assert r["graph_captures"] == 0 and r["seq_fallbacks"] == 0, r
    e.close()
