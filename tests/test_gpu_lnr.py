"""Bad-data identification by the largest normalized residual, on the GPU
against the oracle.

rn_i = |r_i| / sqrt(Omega_ii) with Omega = R - H G^-1 H^T at the estimate
(Abur & Exposito 2004, ch. 5).  The engine forms Omega_ii from the sparse
inverse subset of the gain on its supernodal fronts (root to leaves); the
oracle from column recurrences on its own left-looking factor (pinned against
the dense covariance in tests/test_oracle_pins.py).  The reference only tests
|r|/sigma (sle_engine.cu:959-977) and never fills SLE_Result.largest_residual*
(wls_solver.cu:361-365).
"""
import ctypes as C

import numpy as np
import pytest

from oracle.oracle import Oracle
from state_estimation_with_cuda_amd import gridgen as gg
from tests.helpers import GRIDS, engine_for
from tests.test_oracle_pins import variant14

pytestmark = pytest.mark.gpu

CASES = {"grid14": GRIDS["grid14"], "variant14": variant14, "geo2k": GRIDS["geo2k"], "geo10k": GRIDS["geo10k"]}


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("fp64", [True, False])
def test_normalized_residuals_match_oracle(name, fp64):
    g = CASES[name]()
    e = engine_for(g, double_precision=fp64)
    r = e.solve(mode=1, max_iterations=50, tolerance=1e-10, flat_start=True)
    assert r["status"] == 0, r
    rn, om = e.normalized_residuals()
    o = Oracle(g)
    o.set_state(*e.state())  # the same estimate
    orn, oom = o.normalized_residuals()
    R = g.sigma ** 2
    d_om = float(np.max(np.abs(om - oom) / R))
    d_rn = float(np.max(np.abs(rn - orn)))
    print(f"[lnr {name} fp64={fp64}] max rn={rn.max():.3f} d_omega/sigma^2={d_om:.2e} d_rn={d_rn:.2e} "
          f"critical={int(np.sum((oom <= 1e-10 * R)))}", flush=True)
    assert d_om < 1e-9, d_om
    assert d_rn < 1e-7 * max(1.0, orn.max()), d_rn
    # sum of leverages = number of states
    assert abs(np.sum(1.0 - om / R) - o.ns) < 1e-6 * o.ns
    # the factor left behind serves REALTIME updates
    idx = np.arange(0, g.m, 5, dtype=np.int32)
    e.set_telemetry_indexed(idx, g.z[idx] + 0.1 * g.sigma[idx])
    r = e.solve(mode=0, max_iterations=20, tolerance=1e-8)
    assert r["status"] == 0 and r["refactors"] == 0, r
    e.close()


def _corrupted(seed_row=1234):
    g = GRIDS["geo2k"]()
    bad = seed_row % g.m
    g.z = g.z.copy()
    g.z[bad] += 20 * g.sigma[bad]
    return g, bad


def test_gross_error_identified_and_eliminated():
    """A +20 sigma error: sle_IdentifyBadData (threshold 6) names it, and with
    auto_bad_data_rejection the solve sets it aside and re-estimates: the state
    equals the oracle's without that measurement, and the result reports it."""
    g, bad = _corrupted()
    e = engine_for(g, double_precision=True)
    assert e.solve(mode=1, max_iterations=50, tolerance=1e-10, flat_start=True)["status"] == 0
    rn, _ = e.normalized_residuals()
    assert int(np.argmax(rn)) == bad and rn[bad] > 10, (np.argmax(rn), rn[bad])
    found = e.identify_bad_data(6.0)
    assert list(found) == [bad], found
    # automatic elimination through sle_Solve (SLE_Result fields)
    e.set_config(auto_bad_data_rejection=1, bad_data_threshold=6.0, convergence_tolerance=1e-10)
    res = e.S["Result"]()
    rc = e.L.sle_Solve(e.h, 1, C.byref(res))
    print(f"[lnr auto] rc={rc} status={res.status} iters={res.iterations} bad={res.bad_data_count} "
          f"largest={res.largest_residual:.3f} idx={res.largest_residual_idx}", flush=True)
    assert rc == 0 and res.status == 0
    assert res.bad_data_count == 1 and res.largest_residual < 6.0
    keep = np.arange(g.m) != bad
    g2 = gg.Grid(**{**g.__dict__})
    for f in ("m_type", "m_loc", "m_end", "sigma", "z", "z_true"):
        setattr(g2, f, getattr(g, f)[keep])
    o = Oracle(g2)
    assert o.solve(mode=1, max_iter=50, tol=1e-11, flat_start=True)["status"] == 0
    oV, oth = o.state()
    V, th = e.state()
    assert np.abs(V - oV).max() < 1e-6 and np.abs(th - oth).max() < 1e-6
    # the largest remaining normalized residual is the oracle's on the cleaned set
    orn, _ = o.normalized_residuals()
    assert abs(res.largest_residual - orn.max()) < 1e-6 * orn.max()
    e.close()


def test_measurement_set_aside_and_back():
    """sle_ext_SetMeasurementActive: inactive rows give the oracle's estimate
    without them; re-activated, the original estimate returns."""
    g = GRIDS["geo2k"]()
    e = engine_for(g, double_precision=True)
    assert e.solve(mode=1, max_iterations=50, tolerance=1e-10, flat_start=True)["status"] == 0
    V0, th0 = e.state()
    drop = np.arange(3, g.m, 97)
    for i in drop:
        e.set_measurement_active(int(i), 0)
    assert e.solve(mode=1, max_iterations=50, tolerance=1e-10)["status"] == 0
    keep = np.ones(g.m, bool)
    keep[drop] = False
    g2 = gg.Grid(**{**g.__dict__})
    for f in ("m_type", "m_loc", "m_end", "sigma", "z", "z_true"):
        setattr(g2, f, getattr(g, f)[keep])
    o = Oracle(g2)
    assert o.solve(mode=1, max_iter=50, tol=1e-11, flat_start=True)["status"] == 0
    oV, oth = o.state()
    V, th = e.state()
    assert np.abs(V - oV).max() < 1e-6 and np.abs(th - oth).max() < 1e-6
    rn, om = e.normalized_residuals()
    assert np.all(rn[drop] == 0) and np.all(om[drop] == 0)
    for i in drop:
        e.set_measurement_active(int(i), 1)
    assert e.solve(mode=1, max_iterations=50, tolerance=1e-10)["status"] == 0
    V, th = e.state()
    assert np.abs(V - V0).max() < 1e-8 and np.abs(th - th0).max() < 1e-8
    e.close()
